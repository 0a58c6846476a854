// pf_fast_launch.hpp -- launchers of the compile-time specialised kernels.
// Each returns false when the line length has no specialisation or the tile
// constraints do not hold; the caller then uses the generic kernels.
#pragma once

#include "pf_common.hpp"
#include "pf_fft.hpp"

namespace pfamd {

struct LastAxis;
struct DctGeom;

struct FastCtx
{
  hipStream_t stream;
  size_t lds_budget;
  int ldiv = 1; // y and x passes: tiles of L/ldiv lines (1 or 2)
  // x pass of the generation: output offset of each realization (nullptr:
  // realization b at b * rows * n0)
  const int64_t* ooff = nullptr;
  bool query = false; // report whether a specialised kernel applies; launch nothing
};

template <typename T>
bool fast_gen_last(const FastCtx& c, int N, const GenGeom& g, const LastAxis& a, const T* sq,
                   const double* noise, uint64_t seed, uint64_t first, cplx<T>* B1,
                   const cplx<T>* tw, int batch);
template <typename T>
bool fast_mid(const FastCtx& c, int N, bool inv, const cplx<T>* in, cplx<T>* out, int H, int N0h,
              int nz, int in_rows, int n_in, int out_rows, int n_out, const cplx<T>* tw,
              int batch);
template <typename T>
bool fast_c2r(const FastCtx& c, int M, const cplx<T>* in, T* out, int64_t rows, int H, int n0,
              const cplx<T>* tw, int transform, T param, int batch);
template <typename T>
bool fast_r2c(const FastCtx& c, int M, const T* in, cplx<T>* out, int64_t rows, int H, int n0,
              const cplx<T>* tw, int batch);
// Eigenvalues of a 2D R2C embedding from the x half spectrum of its even
// part (rows 0 .. M of N1 = 2M): y pass, scaling, classification and roots
// (kf_eig_c2r_y; sq of type T). *blocks: classification partials per
// realization (set also in query mode).
template <typename T>
bool fast_eig_c2r_y(const FastCtx& c, int M, const cplx<double>* spec, int Hp, int H, int N0h,
                    double inv_next, double thr0, const double* thrs, double* eig, T* sq,
                    ClassifyPartial* partial, const cplx<double>* tw, int batch, int* blocks);
template <typename T>
bool fast_fused(const FastCtx& c, int N, cplx<T>* C, const GenGeom& g, const LastAxis& a,
                const double* eig, int op, double inv_next, const cplx<T>* tw, int batch,
                const double* spec, int nbins);

// DCT/DST field passes (line half-length M along the pass axis); each
// returns false when M has no specialisation or its tile does not fit.
template <typename T>
bool fast_dct_gen_last(const FastCtx& c, int M, const DctGeom& g, const T* sq,
                       const double* noise, uint64_t seed, uint64_t first, T* G,
                       const cplx<T>* tw, int batch);
template <typename T>
bool fast_dct_mid(const FastCtx& c, int M, const DctGeom& g, const T* G, T* Hh,
                  const cplx<T>* tw, int batch);
template <typename T>
bool fast_dct_x(const FastCtx& c, int M, const DctGeom& g, const T* Hh, T* out, int64_t rows,
                const cplx<T>* tw, int transform, T param, int batch);

} // namespace pfamd
