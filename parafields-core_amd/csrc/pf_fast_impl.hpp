// pf_fast_impl.hpp -- size dispatch of the specialised kernels (included by
// one translation unit per precision).
#pragma once

#include "pf_dctdst_fast.hpp"
#include "pf_fast.hpp"
#include "pf_fast_launch.hpp"

namespace pfamd {

// Lines per tile: the largest power of two with L*N complex values in 64 KB
// (FP32: 16 lines x 512; FP64: 8 x 512), one thread per 16 values.
// The complex/real "H" arguments of the launchers below are row pitches.
template <typename T>
constexpr int fast_L(int N)
{
  int L = sizeof(T) == 4 ? 16 : 8;
  const int lim = sizeof(T) == 4 ? 8192 : 4096;
  while (L > 1 && L * N > lim)
    L /= 2;
  return L;
}

// Lines per tile of the x passes, whose LDS tile is padded to L + 1 lines:
// the largest power of two <= fast_L with (L + 1) * M complex values in 64 KB.
template <typename T>
constexpr int pad_L(int M)
{
  int L = fast_L<T>(M);
  while (L > 1 && (L + 1) * M * 2 * (int)sizeof(T) > 65536)
    L /= 2;
  return L;
}

template <typename T>
constexpr int fast_NT(int N, int L)
{
  // values per thread and stage: 16 for FP32 (512 threads for a 16 x 512
  // tile), 8 for FP64, whose values take twice the registers (more waves in
  // flight per tile)
  const int me = sizeof(T) == 4 ? 16 : 8;
  int nt = (L * N + me - 1) / me;
  nt = ((nt + 63) / 64) * 64;
  return nt < 64 ? 64 : nt;
}

#define PF_FAST_SIZES(X)                                                                           \
  X(16) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536) X(2048)

template <typename T, int N>
bool gen_last_n(const FastCtx& c, const GenGeom& g, const LastAxis& a, const T* sq,
                const double* noise, uint64_t seed, uint64_t first, cplx<T>* B1,
                const cplx<T>* tw, int batch)
{
  constexpr int L = fast_L<T>(N), NT = fast_NT<T>(N, L);
  const size_t lds = (size_t)L * N * sizeof(cplx<T>);
  if (a.N0h % L != 0 || lds > c.lds_budget)
    return false;
  if (c.query)
    return true;
  dim3 grid(a.N0h / L + 1, a.Nmid, batch);
  // 32-bit slot walk when every storage index and eigenvalue offset of a
  // realization stays below 2^31 and its eigenvalue table (at most one entry
  // per half-spectrum entry) and B1 slab stay below 4 GiB
  const int64_t entries = (int64_t)a.H * a.Nmid * N;
  const int64_t lim = (int64_t)1 << 32;
  const bool idx64 = entries >= ((int64_t)1 << 31) || g.noise_stride >= lim ||
                     entries * (int64_t)sizeof(T) >= lim ||
                     (int64_t)a.Hp * a.Nmid * a.nlast * (int64_t)sizeof(cplx<T>) >= lim;
  // outputs j < nlast <= N/2 only: the last stage skips the upper half
  const bool half = 2 * a.nlast <= N;
  if (noise)
    kf_gen_last<T, N, L, NT, 0, true><<<grid, NT, lds, c.stream>>>(g, a, sq, noise, seed, first,
                                                                   B1, tw);
  else if (idx64)
    kf_gen_last<T, N, L, NT, 1, true><<<grid, NT, lds, c.stream>>>(g, a, sq, noise, seed, first,
                                                                   B1, tw);
  else if (half)
    kf_gen_last<T, N, L, NT, 1, false, true><<<grid, NT, lds, c.stream>>>(g, a, sq, noise, seed,
                                                                          first, B1, tw);
  else
    kf_gen_last<T, N, L, NT, 1, false><<<grid, NT, lds, c.stream>>>(g, a, sq, noise, seed, first,
                                                                    B1, tw);
  return true;
}

template <typename T>
bool fast_gen_last(const FastCtx& c, int N, const GenGeom& g, const LastAxis& a, const T* sq,
                   const double* noise, uint64_t seed, uint64_t first, cplx<T>* B1,
                   const cplx<T>* tw, int batch)
{
  switch (N) {
#define X(n)                                                                                       \
  case n:                                                                                          \
    return gen_last_n<T, n>(c, g, a, sq, noise, seed, first, B1, tw, batch);
    PF_FAST_SIZES(X)
#undef X
    default:
      return false;
  }
}

template <typename T, int N, int L>
bool mid_nl(const FastCtx& c, bool inv, const cplx<T>* in, cplx<T>* out, int H, int N0h, int nz,
            int in_rows, int n_in, int out_rows, int n_out, const cplx<T>* tw, int batch)
{
  constexpr int NT = fast_NT<T>(N, L);
  const size_t lds = (size_t)L * N * sizeof(cplx<T>);
  // byte offsets within a tile (rows of N elements, Nyquist-column groups
  // of L row blocks) fit in 32 bits
  const int64_t rows = in_rows > out_rows ? in_rows : out_rows;
  if (N0h % L != 0 || lds > c.lds_budget ||
      ((int64_t)N + (int64_t)L * rows) * H * (int64_t)sizeof(cplx<T>) >= ((int64_t)1 << 32))
    return false;
  if (c.query)
    return true;
  dim3 grid(N0h / L + 1, nz, batch);
  if (inv && 2 * n_out <= N)
    kf_mid<T, N, L, NT, true, true><<<grid, NT, lds, c.stream>>>(in, out, H, N0h, nz, in_rows,
                                                                 n_in, out_rows, n_out, tw);
  else if (inv)
    kf_mid<T, N, L, NT, true><<<grid, NT, lds, c.stream>>>(in, out, H, N0h, nz, in_rows, n_in,
                                                           out_rows, n_out, tw);
  else
    kf_mid<T, N, L, NT, false><<<grid, NT, lds, c.stream>>>(in, out, H, N0h, nz, in_rows, n_in,
                                                            out_rows, n_out, tw);
  return true;
}

// Half-width tiles (L/2 lines, half the LDS) let a CU hold more blocks next to
// the generation pass; the two tile halves of a 128-byte row segment run on
// one XCD (neighbouring kx tiles, x-fastest block order).
template <typename T, int N>
bool mid_n(const FastCtx& c, bool inv, const cplx<T>* in, cplx<T>* out, int H, int N0h, int nz,
           int in_rows, int n_in, int out_rows, int n_out, const cplx<T>* tw, int batch)
{
  constexpr int L = fast_L<T>(N);
  if constexpr (L >= 8) {
    if (c.ldiv == 2)
      return mid_nl<T, N, L / 2>(c, inv, in, out, H, N0h, nz, in_rows, n_in, out_rows, n_out, tw,
                                 batch);
  }
  return mid_nl<T, N, L>(c, inv, in, out, H, N0h, nz, in_rows, n_in, out_rows, n_out, tw, batch);
}

template <typename T>
bool fast_mid(const FastCtx& c, int N, bool inv, const cplx<T>* in, cplx<T>* out, int H, int N0h,
              int nz, int in_rows, int n_in, int out_rows, int n_out, const cplx<T>* tw,
              int batch)
{
  switch (N) {
#define X(n)                                                                                       \
  case n:                                                                                          \
    return mid_n<T, n>(c, inv, in, out, H, N0h, nz, in_rows, n_in, out_rows, n_out, tw, batch);
    PF_FAST_SIZES(X)
#undef X
    default:
      return false;
  }
}

template <typename T, int M, int L>
bool c2r_nl(const FastCtx& c, const cplx<T>* in, T* out, int64_t rows, int H, int n0,
            const cplx<T>* tw, int transform, T param, int batch)
{
  constexpr int NT = fast_NT<T>(M, L);
  const size_t lds = (size_t)(L + 1) * M * sizeof(cplx<T>);
  if (lds > c.lds_budget)
    return false;
  if (c.query)
    return true;
  dim3 grid((unsigned)((rows + L - 1) / L), batch);
  kf_c2r_x<T, T, M, L, NT><<<grid, NT, lds, c.stream>>>(in, out, rows, H, n0, tw, transform, param,
                                                        c.ooff);
  return true;
}

template <typename T, int M>
bool c2r_n(const FastCtx& c, const cplx<T>* in, T* out, int64_t rows, int H, int n0,
           const cplx<T>* tw, int transform, T param, int batch)
{
  constexpr int L = pad_L<T>(M);
  if constexpr (L >= 8) {
    if (c.ldiv == 2)
      return c2r_nl<T, M, L / 2>(c, in, out, rows, H, n0, tw, transform, param, batch);
  }
  return c2r_nl<T, M, L>(c, in, out, rows, H, n0, tw, transform, param, batch);
}

template <typename T>
bool fast_c2r(const FastCtx& c, int M, const cplx<T>* in, T* out, int64_t rows, int H, int n0,
              const cplx<T>* tw, int transform, T param, int batch)
{
  switch (M) {
#define X(n)                                                                                       \
  case n:                                                                                          \
    return c2r_n<T, n>(c, in, out, rows, H, n0, tw, transform, param, batch);
    PF_FAST_SIZES(X)
#undef X
    default:
      return false;
  }
}

template <typename T, int M>
bool r2c_n(const FastCtx& c, const T* in, cplx<T>* out, int64_t rows, int H, int n0,
           const cplx<T>* tw, int batch)
{
  constexpr int L = pad_L<T>(M), NT = fast_NT<T>(M, L);
  const size_t lds = (size_t)(L + 1) * M * sizeof(cplx<T>);
  if (lds > c.lds_budget)
    return false;
  if (c.query)
    return true;
  dim3 grid((unsigned)((rows + L - 1) / L), batch);
  kf_r2c_x<T, T, M, L, NT><<<grid, NT, lds, c.stream>>>(in, out, rows, H, n0, tw);
  return true;
}

template <typename T>
bool fast_r2c(const FastCtx& c, int M, const T* in, cplx<T>* out, int64_t rows, int H, int n0,
              const cplx<T>* tw, int batch)
{
  switch (M) {
#define X(n)                                                                                       \
  case n:                                                                                          \
    return r2c_n<T, n>(c, in, out, rows, H, n0, tw, batch);
    PF_FAST_SIZES(X)
#undef X
    default:
      return false;
  }
}

template <typename TS, int M>
bool eig_c2r_y_n(const FastCtx& c, const cplx<double>* spec, int Hp, int H, int N0h,
                 double inv_next, double thr0, const double* thrs, double* eig, TS* sq,
                 ClassifyPartial* partial, const cplx<double>* tw, int batch, int* blocks)
{
  // full-width tiles: 4 lines at M = 1024 (64 KB); half-width tiles (more
  // blocks per CU, 32-byte row segments) measured slower
  constexpr int L = fast_L<double>(M);
  constexpr int NT = fast_NT<double>(M, L);
  const size_t lds = (size_t)L * M * sizeof(cplx<double>);
  // 32-bit offsets: input rows within one realization's span, outputs
  // within one realization's [2M][H] tables
  if (N0h % L != 0 || lds > c.lds_budget || H > Hp ||
      (int64_t)(M + 1) * Hp * (int64_t)sizeof(cplx<double>) >= ((int64_t)1 << 32) ||
      (int64_t)2 * M * H >= ((int64_t)1 << 31))
    return false;
  const int nb = N0h / L + 1;
  if (blocks)
    *blocks = nb;
  if (c.query)
    return true;
  kf_eig_c2r_y<TS, M, L, NT><<<dim3(nb, batch), NT, lds, c.stream>>>(
    spec, Hp, H, N0h, inv_next, thr0, thrs, eig, sq, partial, tw);
  return true;
}

template <typename TS>
bool fast_eig_c2r_y(const FastCtx& c, int M, const cplx<double>* spec, int Hp, int H, int N0h,
                    double inv_next, double thr0, const double* thrs, double* eig, TS* sq,
                    ClassifyPartial* partial, const cplx<double>* tw, int batch, int* blocks)
{
  switch (M) {
#define X(n)                                                                                       \
  case n:                                                                                          \
    return eig_c2r_y_n<TS, n>(c, spec, Hp, H, N0h, inv_next, thr0, thrs, eig, sq, partial, tw,    \
                              batch, blocks);
    PF_FAST_SIZES(X)
#undef X
    default:
      return false;
  }
}

// The fused pass runs two transforms per tile: 8 values per thread and stage
// (twice the threads of the single-transform passes) keep more waves in
// flight through its LDS stages.
template <typename T, int N>
bool fused_n(const FastCtx& c, cplx<T>* C, const GenGeom& g, const LastAxis& a, const double* eig,
             int op, double inv_next, const cplx<T>* tw, int batch, const double* spec, int nbins)
{
  constexpr int L = fast_L<T>(N);
  constexpr int NT = (L * N / 8 + 63) / 64 * 64 < 64 ? 64 : (L * N / 8 + 63) / 64 * 64;
  const size_t lds = (size_t)L * N * sizeof(cplx<T>);
  if (a.N0h % L != 0 || lds > c.lds_budget)
    return false;
  // 32-bit element offsets within one realization's span
  if ((int64_t)a.Hp * a.Nmid * a.nlast * (int64_t)sizeof(cplx<T>) >= ((int64_t)1 << 32))
    return false;
  if (c.query)
    return true;
  dim3 grid(a.N0h / L + 1, a.Nmid, batch);
  if (2 * a.nlast <= N)
    kf_fused_mult<T, N, L, NT, true><<<grid, NT, lds, c.stream>>>(C, g, a, eig, op, inv_next, tw,
                                                                  spec, nbins);
  else
    kf_fused_mult<T, N, L, NT, false><<<grid, NT, lds, c.stream>>>(C, g, a, eig, op, inv_next,
                                                                   tw, spec, nbins);
  return true;
}

template <typename T>
bool fast_fused(const FastCtx& c, int N, cplx<T>* C, const GenGeom& g, const LastAxis& a,
                const double* eig, int op, double inv_next, const cplx<T>* tw, int batch,
                const double* spec, int nbins)
{
  switch (N) {
#define X(n)                                                                                       \
  case n:                                                                                          \
    return fused_n<T, n>(c, C, g, a, eig, op, inv_next, tw, batch, spec, nbins);
    PF_FAST_SIZES(X)
#undef X
    default:
      return false;
  }
}

// --- DCT/DST field passes ---------------------------------------------------
// Tiles of L lines of M complex values (no pad line, X_M folded into row 0):
// the largest power of two L <= 32 (FP32) / 16 (FP64) whose tile fits in
// 64 KB (FP32: 8 x 1024, 16 x 512, 32 x 256).
template <typename T>
constexpr int dct_L(int M)
{
  int L = sizeof(T) == 4 ? 32 : 16;
  while (L > 1 && (size_t)L * M * 2 * sizeof(T) > (size_t)64 * 1024)
    L /= 2;
  return L;
}

template <typename T, int M>
constexpr size_t dct_lds()
{
  return (size_t)dct_L<T>(M) * M * 2 * sizeof(T);
}

// Threads of a DCT/DST pass: 16 values per thread for FP32, 8 for FP64 (its
// values take twice the registers).
template <typename T>
constexpr int dct_NT(int M, int L)
{
  const int me = sizeof(T) == 4 ? 16 : 8;
  int nt = (L * M + me - 1) / me;
  nt = ((nt + 63) / 64) * 64;
  return nt < 64 ? 64 : nt;
}

#define PF_DCT_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024)

template <typename T, int M>
bool dct_gen_last_n(const FastCtx& c, const DctGeom& g, const T* sq, const double* noise,
                    uint64_t seed, uint64_t first, T* G, const cplx<T>* tw, int batch)
{
  constexpr int L = dct_L<T>(M), NT = dct_NT<T>(M, L);
  constexpr size_t lds = dct_lds<T, M>();
  const int Pm = g.dim == 3 ? g.P1 : 1;
  const int64_t nlast = g.dim == 3 ? g.n2 : g.n1;
  // the pass stores at 32-bit element offsets within one output array
  if ((g.P0 - 1) % L != 0 || lds > c.lds_budget ||
      nlast * Pm * (int64_t)g.P0p >= ((int64_t)1 << 32))
    return false;
  if (c.query)
    return true;
  dim3 grid((g.P0 - 1) / L + 1, Pm, batch);
  if (noise)
    kdf_gen_last<T, M, L, NT, 0><<<grid, NT, lds, c.stream>>>(g, sq, noise, seed, first, G, tw);
  else
    kdf_gen_last<T, M, L, NT, 1><<<grid, NT, lds, c.stream>>>(g, sq, noise, seed, first, G, tw);
  return true;
}

template <typename T>
bool fast_dct_gen_last(const FastCtx& c, int M, const DctGeom& g, const T* sq,
                       const double* noise, uint64_t seed, uint64_t first, T* G,
                       const cplx<T>* tw, int batch)
{
  switch (M) {
#define X(n)                                                                                       \
  case n:                                                                                          \
    return dct_gen_last_n<T, n>(c, g, sq, noise, seed, first, G, tw, batch);
    PF_DCT_SIZES(X)
#undef X
    default:
      return false;
  }
}

template <typename T, int M>
bool dct_mid_n(const FastCtx& c, const DctGeom& g, const T* G, T* Hh, const cplx<T>* tw,
               int batch)
{
  constexpr int L = dct_L<T>(M), NT = dct_NT<T>(M, L);
  constexpr size_t lds = dct_lds<T, M>();
  if ((g.P0 - 1) % L != 0 || lds > c.lds_budget)
    return false;
  if (c.query)
    return true;
  dim3 grid((g.P0 - 1) / L + 1, g.n2, batch);
  kdf_mid<T, M, L, NT><<<grid, NT, lds, c.stream>>>(g, G, Hh, tw);
  return true;
}

template <typename T>
bool fast_dct_mid(const FastCtx& c, int M, const DctGeom& g, const T* G, T* Hh,
                  const cplx<T>* tw, int batch)
{
  switch (M) {
#define X(n)                                                                                       \
  case n:                                                                                          \
    return dct_mid_n<T, n>(c, g, G, Hh, tw, batch);
    PF_DCT_SIZES(X)
#undef X
    default:
      return false;
  }
}

template <typename T, int M>
bool dct_x_n(const FastCtx& c, const DctGeom& g, const T* Hh, T* out, int64_t rows,
             const cplx<T>* tw, int transform, T param, int batch)
{
  constexpr int L = dct_L<T>(M), NT = dct_NT<T>(M, L);
  constexpr size_t lds = dct_lds<T, M>();
  if (lds > c.lds_budget)
    return false;
  if (c.query)
    return true;
  dim3 grid((unsigned)((rows + L - 1) / L), batch);
  kdf_x<T, M, L, NT><<<grid, NT, lds, c.stream>>>(g, Hh, out, rows, tw, transform, param);
  return true;
}

template <typename T>
bool fast_dct_x(const FastCtx& c, int M, const DctGeom& g, const T* Hh, T* out, int64_t rows,
                const cplx<T>* tw, int transform, T param, int batch)
{
  switch (M) {
#define X(n)                                                                                       \
  case n:                                                                                          \
    return dct_x_n<T, n>(c, g, Hh, out, rows, tw, transform, param, batch);
    PF_DCT_SIZES(X)
#undef X
    default:
      return false;
  }
}

#define PF_INSTANTIATE(T)                                                                          \
  template bool fast_gen_last<T>(const FastCtx&, int, const GenGeom&, const LastAxis&, const T*,   \
                                 const double*, uint64_t, uint64_t, cplx<T>*, const cplx<T>*,      \
                                 int);                                                             \
  template bool fast_mid<T>(const FastCtx&, int, bool, const cplx<T>*, cplx<T>*, int, int, int,    \
                            int, int, int, int, const cplx<T>*, int);                              \
  template bool fast_c2r<T>(const FastCtx&, int, const cplx<T>*, T*, int64_t, int, int,            \
                            const cplx<T>*, int, T, int);                                          \
  template bool fast_r2c<T>(const FastCtx&, int, const T*, cplx<T>*, int64_t, int, int,            \
                            const cplx<T>*, int);                                                  \
  template bool fast_eig_c2r_y<T>(const FastCtx&, int, const cplx<double>*, int, int, int, double,  \
                                  double, const double*, double*, T*, ClassifyPartial*,            \
                                  const cplx<double>*, int, int*);                                 \
  template bool fast_fused<T>(const FastCtx&, int, cplx<T>*, const GenGeom&, const LastAxis&,      \
                              const double*, int, double, const cplx<T>*, int, const double*,      \
                              int);                                                                \
  template bool fast_dct_gen_last<T>(const FastCtx&, int, const DctGeom&, const T*, const double*, \
                                     uint64_t, uint64_t, T*, const cplx<T>*, int);                 \
  template bool fast_dct_mid<T>(const FastCtx&, int, const DctGeom&, const T*, T*,                 \
                                const cplx<T>*, int);                                              \
  template bool fast_dct_x<T>(const FastCtx&, int, const DctGeom&, const T*, T*, int64_t,          \
                              const cplx<T>*, int, T, int);

} // namespace pfamd
