// pf_fft.hpp -- LDS-resident Stockham FFT for a tile of lines (gfx950).
//
// A workgroup holds L lines of length N in LDS. Element e of line l lives at
// buf[e*se + l*sl]: element-major tiles (se = L, sl = 1) serve the strided
// axes, whose lines are loaded coalesced across l; line-major tiles (se = 1,
// sl = N) serve the contiguous x axis. Each Stockham stage (radix R) reads R
// elements N/R apart, applies the twiddles exp(-+2 pi i r k / (Ns R)), does
// the radix-R butterfly in registers and writes the self-sorted output
// (Govindaraju et al., SC'08 formulation); every thread keeps at most MAXE
// elements in registers per stage.
#pragma once

#include "pf_device.hpp"

namespace pfamd {

constexpr int kMaxStages = 12;

struct Sched
{
  int n;
  int nst;
  int rad[kMaxStages];
};

// e^{-2 pi i j / R} = (c[j], -s[j]) with c, s the cosine and sine of 2 pi j / R
template <int R>
struct Roots;
template <>
struct Roots<3>
{
  static constexpr double c[3] = { 1.0, -0.5, -0.5 };
  static constexpr double s[3] = { 0.0, 0.86602540378443864676, -0.86602540378443864676 };
};
template <>
struct Roots<5>
{
  static constexpr double c[5] = { 1.0, 0.30901699437494742410, -0.80901699437494742410,
                                   -0.80901699437494742410, 0.30901699437494742410 };
  static constexpr double s[5] = { 0.0, 0.95105651629515357212, 0.58778525229247312917,
                                   -0.58778525229247312917, -0.95105651629515357212 };
};
template <>
struct Roots<7>
{
  static constexpr double c[7] = { 1.0,
                                   0.62348980185873353053,
                                   -0.22252093395631440429,
                                   -0.90096886790241912624,
                                   -0.90096886790241912624,
                                   -0.22252093395631440429,
                                   0.62348980185873353053 };
  static constexpr double s[7] = { 0.0,
                                   0.78183148246802980871,
                                   0.97492791218182360702,
                                   0.43388373911755812048,
                                   -0.43388373911755812048,
                                   -0.97492791218182360702,
                                   -0.78183148246802980871 };
};

template <typename T, int R, bool INV>
struct Butterfly
{
  // generic odd radix: O(R^2) with constant roots
  static __device__ __forceinline__ void run(cplx<T>* a)
  {
    cplx<T> y[R];
#pragma unroll
    for (int s = 0; s < R; s++) {
      cplx<T> acc = a[0];
#pragma unroll
      for (int q = 1; q < R; q++) {
        const int j = (q * s) % R;
        const T c = (T)Roots<R>::c[j];
        const T sn = INV ? (T)Roots<R>::s[j] : -(T)Roots<R>::s[j];
        acc.x += a[q].x * c - a[q].y * sn;
        acc.y += a[q].x * sn + a[q].y * c;
      }
      y[s] = acc;
    }
#pragma unroll
    for (int s = 0; s < R; s++)
      a[s] = y[s];
  }
};

template <typename T, bool INV>
struct Butterfly<T, 2, INV>
{
  static __device__ __forceinline__ void run(cplx<T>* a)
  {
    const cplx<T> t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <typename T, bool INV>
__device__ __forceinline__ void radix4(cplx<T>& a0, cplx<T>& a1, cplx<T>& a2, cplx<T>& a3)
{
  const cplx<T> t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3);
  const cplx<T> d = csub(a1, a3);
  // forward: (a1 - a3) * (-i); inverse: * (+i)
  const cplx<T> t3 = INV ? cplx<T>{ -d.y, d.x } : cplx<T>{ d.y, -d.x };
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <typename T, bool INV>
struct Butterfly<T, 4, INV>
{
  static __device__ __forceinline__ void run(cplx<T>* a) { radix4<T, INV>(a[0], a[1], a[2], a[3]); }
};

template <typename T, bool INV>
struct Butterfly<T, 8, INV>
{
  static __device__ __forceinline__ void run(cplx<T>* a)
  {
    // decimation in time: even and odd radix-4 DFTs, then W8^k combine
    cplx<T> e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
    cplx<T> o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
    radix4<T, INV>(e0, e1, e2, e3);
    radix4<T, INV>(o0, o1, o2, o3);
    const T r = (T)0.70710678118654752440;
    // W8^1 = (1 -+ i)/sqrt2, W8^2 = -+i, W8^3 = (-1 -+ i)/sqrt2
    cplx<T> w1o1, w2o2, w3o3;
    if (!INV) {
      w1o1 = { (o1.x + o1.y) * r, (o1.y - o1.x) * r };
      w2o2 = { o2.y, -o2.x };
      w3o3 = { (o3.y - o3.x) * r, -(o3.x + o3.y) * r };
    } else {
      w1o1 = { (o1.x - o1.y) * r, (o1.y + o1.x) * r };
      w2o2 = { -o2.y, o2.x };
      w3o3 = { -(o3.x + o3.y) * r, (o3.x - o3.y) * r };
    }
    a[0] = cadd(e0, o0);
    a[4] = csub(e0, o0);
    a[1] = cadd(e1, w1o1);
    a[5] = csub(e1, w1o1);
    a[2] = cadd(e2, w2o2);
    a[6] = csub(e2, w2o2);
    a[3] = cadd(e3, w3o3);
    a[7] = csub(e3, w3o3);
  }
};

// One Stockham stage over the L lines held in LDS. tw[t] = e^{-2 pi i t / Ntab},
// Ntab = N * twmul. Ends with a barrier.
template <typename T, int R, int MAXE, bool LM, bool INV>
__device__ __forceinline__ void fft_stage(cplx<T>* buf, const cplx<T>* tw, int twmul, int N, int L,
                                          int Ns, int se, int sl)
{
  constexpr int IT = (MAXE + R - 1) / R;
  const int NR = N / R;
  const int total = L * NR;
  const int tstep = (N / (Ns * R)) * twmul;
  cplx<T> v[IT][R];
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int w = threadIdx.x + it * blockDim.x;
    if (w < total) {
      int l, j;
      if (LM) {
        j = w % NR;
        l = w / NR;
      } else {
        l = w % L;
        j = w / L;
      }
      const int k = j % Ns;
#pragma unroll
      for (int r = 0; r < R; r++)
        v[it][r] = buf[(j + r * NR) * se + l * sl];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) {
          cplx<T> t = tw[r * k * tstep];
          if (INV)
            t.y = -t.y;
          v[it][r] = cmul(v[it][r], t);
        }
      }
      Butterfly<T, R, INV>::run(v[it]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int w = threadIdx.x + it * blockDim.x;
    if (w < total) {
      int l, j;
      if (LM) {
        j = w % NR;
        l = w / NR;
      } else {
        l = w % L;
        j = w / L;
      }
      const int k = j % Ns;
      const int base = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; r++)
        buf[(base + r * Ns) * se + l * sl] = v[it][r];
    }
  }
  __syncthreads();
}

// Full transform of the L lines in LDS (unnormalised; INV = e^{+i}).
template <typename T, int MAXE, bool LM, bool INV>
__device__ void fft_lds(cplx<T>* buf, const cplx<T>* tw, int twmul, const Sched& s, int L, int se,
                        int sl)
{
  int Ns = 1;
  for (int st = 0; st < s.nst; st++) {
    const int R = s.rad[st];
    switch (R) {
      case 8:
        fft_stage<T, 8, MAXE, LM, INV>(buf, tw, twmul, s.n, L, Ns, se, sl);
        break;
      case 4:
        fft_stage<T, 4, MAXE, LM, INV>(buf, tw, twmul, s.n, L, Ns, se, sl);
        break;
      case 2:
        fft_stage<T, 2, MAXE, LM, INV>(buf, tw, twmul, s.n, L, Ns, se, sl);
        break;
      case 3:
        fft_stage<T, 3, MAXE, LM, INV>(buf, tw, twmul, s.n, L, Ns, se, sl);
        break;
      case 5:
        fft_stage<T, 5, MAXE, LM, INV>(buf, tw, twmul, s.n, L, Ns, se, sl);
        break;
      case 7:
        fft_stage<T, 7, MAXE, LM, INV>(buf, tw, twmul, s.n, L, Ns, se, sl);
        break;
      default:
        break;
    }
    Ns *= R;
  }
}

} // namespace pfamd
