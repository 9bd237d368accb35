// ptc_fft2.hip -- compile-time specialised, register-resident FFT passes.
//
// For the transform lengths of the production grids the radix plan, the
// number of stages and the threads per line are compile-time constants, so
// every butterfly lives in registers and no private (scratch) memory is used.
// Stage 0 reads its inputs straight from HBM and the last stage writes its
// outputs straight to HBM; LDS only carries the exchanges between stages.
// The algorithm is the same self-sorting Stockham transform as the generic
// kernels of ptc_fft.hip (which remain the path for other lengths):
//
//   stage with radix R after a span Ns:  butterfly b reads x[b + r N/R],
//   multiplies by w^(r (b mod Ns) N/(Ns R)), applies the R-point DFT and
//   writes x[(b div Ns) Ns R + (b mod Ns) + r Ns].
//
// Twiddles: w^e for the first power comes from the host FP64 table; the
// higher powers of the same stage are formed by complex products (at most six
// products, error ~ 7 ulp), which replaces R-1 table loads per butterfly.
#include <algorithm>

#include "ptc_internal.hpp"

namespace ptc {

namespace f2 {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
template <bool INV>
__device__ __forceinline__ double2 mul_mi(double2 a) {
  return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// ---- compile-time radix plan: 8s first, then 4 or 2, then 9, 3, 5, 7 ----
struct PlanC {
  int ns;
  int r[12];
};

constexpr PlanC make_plan(int n) {
  PlanC p{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int m = n;
  while (m % 8 == 0) { p.r[p.ns++] = 8; m /= 8; }
  if (m % 4 == 0) { p.r[p.ns++] = 4; m /= 4; }
  if (m % 2 == 0) { p.r[p.ns++] = 2; m /= 2; }
  while (m % 9 == 0) { p.r[p.ns++] = 9; m /= 9; }
  while (m % 3 == 0) { p.r[p.ns++] = 3; m /= 3; }
  while (m % 5 == 0) { p.r[p.ns++] = 5; m /= 5; }
  while (m % 7 == 0) { p.r[p.ns++] = 7; m /= 7; }
  if (m != 1) p.ns = -1;
  return p;
}

constexpr int plan_rmax(int n) {
  PlanC p = make_plan(n);
  int mx = 1;
  for (int s = 0; s < p.ns; ++s) mx = p.r[s] > mx ? p.r[s] : mx;
  return mx;
}

constexpr int plan_span(int n, int s) {  // product of the radices before stage s
  PlanC p = make_plan(n);
  int v = 1;
  for (int q = 0; q < s; ++q) v *= p.r[q];
  return v;
}

// spectra are stored as double2 (FP64 form) or float2 (FP32-storage form);
// arithmetic is FP64 either way
__device__ __forceinline__ double2 ld2(const double2 &v) { return v; }
__device__ __forceinline__ double2 ld2(const float2 &v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ void st2(double2 &d, double2 v) { d = v; }
__device__ __forceinline__ void st2(float2 &d, double2 v) { d = make_float2((float)v.x, (float)v.y); }

// Streaming (nontemporal) forms: the access is marked as not to be kept in
// the caches.  A pass over a field larger than the last-level cache reads
// and writes each value once, so keeping it only evicts other data.
typedef double st_d2 __attribute__((ext_vector_type(2)));
typedef float st_f2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld2s(const double2 &v) {
  const st_d2 w = __builtin_nontemporal_load(reinterpret_cast<const st_d2 *>(&v));
  return make_double2(w.x, w.y);
}
__device__ __forceinline__ double2 ld2s(const float2 &v) {
  const st_f2 w = __builtin_nontemporal_load(reinterpret_cast<const st_f2 *>(&v));
  return make_double2(w.x, w.y);
}
__device__ __forceinline__ void st2s(double2 &d, double2 v) {
  st_d2 w = {v.x, v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<st_d2 *>(&d));
}
__device__ __forceinline__ void st2s(float2 &d, double2 v) {
  st_f2 w = {(float)v.x, (float)v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<st_f2 *>(&d));
}
template <bool STREAM, typename S>
__device__ __forceinline__ double2 ld2x(const S &v) {
  if constexpr (STREAM) return ld2s(v);
  else return ld2(v);
}
template <bool STREAM, typename S>
__device__ __forceinline__ void st2x(S &d, double2 v) {
  if constexpr (STREAM) st2s(d, v);
  else st2(d, v);
}

// LDS padding of one line: one slot per 8 elements
__device__ __forceinline__ int pad8(int m) { return m + (m >> 3); }
constexpr int padded(int n) { return n + n / 8 + 1; }

// ---- small DFTs ----
template <int R, bool INV>
struct Dft;
template <bool INV>
struct Dft<2, INV> {
  static __device__ __forceinline__ void run(double2 *v) {
    const double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};
template <bool INV>
struct Dft<4, INV> {
  static __device__ __forceinline__ void run(double2 *v) {
    const double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const double2 t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[1] = cadd(t1, t3);
    v[2] = csub(t0, t2);
    v[3] = csub(t1, t3);
  }
};
template <bool INV>
struct Dft<8, INV> {
  static __device__ __forceinline__ void run(double2 *v) {
    constexpr double r2 = 0.70710678118654752440;
    double2 e[4] = {v[0], v[2], v[4], v[6]};
    double2 o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, INV>::run(e);
    Dft<4, INV>::run(o);
    const double2 o1 = INV ? make_double2(r2 * (o[1].x - o[1].y), r2 * (o[1].x + o[1].y))
                           : make_double2(r2 * (o[1].x + o[1].y), r2 * (o[1].y - o[1].x));
    const double2 o2 = mul_mi<INV>(o[2]);
    const double2 o3 = INV ? make_double2(-r2 * (o[3].x + o[3].y), r2 * (o[3].x - o[3].y))
                           : make_double2(r2 * (o[3].y - o[3].x), -r2 * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

template <int R>
struct Trig;
template <>
struct Trig<3> {
  static constexpr double c[3] = {1.0, -0.5, -0.5};
  static constexpr double s[3] = {0.0, 0.86602540378443864676, -0.86602540378443864676};
};
template <>
struct Trig<5> {
  static constexpr double c[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410, -0.80901699437494742410,
                                  0.30901699437494742410};
  static constexpr double s[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917, -0.58778525229247312917,
                                  -0.95105651629515357212};
};
template <>
struct Trig<7> {
  static constexpr double c[7] = {1.0, 0.62348980185873353053, -0.22252093395631440429, -0.90096886790241912624,
                                  -0.90096886790241912624, -0.22252093395631440429, 0.62348980185873353053};
  static constexpr double s[7] = {0.0, 0.78183148246802980871, 0.97492791218182360702, 0.43388373911755812048,
                                  -0.43388373911755812048, -0.97492791218182360702, -0.78183148246802980871};
};

template <int R, bool INV>
struct DftOdd {
  static __device__ __forceinline__ void run(double2 *v) {
    constexpr int H = (R - 1) / 2;
    double2 a[H + 1], b[H + 1];
#pragma unroll
    for (int q = 1; q <= H; ++q) {
      a[q] = cadd(v[q], v[R - q]);
      b[q] = csub(v[q], v[R - q]);
    }
    double2 x0 = v[0];
#pragma unroll
    for (int q = 1; q <= H; ++q) x0 = cadd(x0, a[q]);
    double2 out[R];
    out[0] = x0;
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      double2 A = v[0], B = make_double2(0.0, 0.0);
#pragma unroll
      for (int q = 1; q <= H; ++q) {
        const double cq = Trig<R>::c[(q * r) % R], sq = Trig<R>::s[(q * r) % R];
        A.x += a[q].x * cq;
        A.y += a[q].y * cq;
        B.x += b[q].x * sq;
        B.y += b[q].y * sq;
      }
      if (!INV) {
        out[r] = make_double2(A.x + B.y, A.y - B.x);
        out[R - r] = make_double2(A.x - B.y, A.y + B.x);
      } else {
        out[r] = make_double2(A.x - B.y, A.y + B.x);
        out[R - r] = make_double2(A.x + B.y, A.y - B.x);
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = out[r];
  }
};
template <bool INV>
struct Dft<3, INV> : DftOdd<3, INV> {};
template <bool INV>
struct Dft<5, INV> : DftOdd<5, INV> {};
template <bool INV>
struct Dft<7, INV> : DftOdd<7, INV> {};
// 9 points as 3 x 3: three 3-point DFTs of the stride-3 subsequences, the
// twiddles w9^(n2 k1), three 3-point DFTs across them; X[k1 + 3 k2]
template <bool INV>
struct Dft<9, INV> {
  static __device__ __forceinline__ void run(double2 *v) {
    // cos and sin of 2 pi k / 9, k = 1, 2, 4
    constexpr double c1 = 0.7660444431189780352024, s1 = 0.6427876096865393263226;
    constexpr double c2 = 0.1736481776669303488517, s2 = 0.9848077530122080593667;
    constexpr double c4 = -0.9396926207859083840541, s4 = 0.3420201433256687330441;
    double2 t[3][3];
#pragma unroll
    for (int n2 = 0; n2 < 3; ++n2) {
      t[n2][0] = v[n2];
      t[n2][1] = v[n2 + 3];
      t[n2][2] = v[n2 + 6];
      Dft<3, INV>::run(t[n2]);
    }
    // w9^k = (cos, -sin) forward, (cos, sin) inverse
    auto tw = [](double2 a, double c, double sn) {
      const double sg = INV ? sn : -sn;
      return make_double2(a.x * c - a.y * sg, a.x * sg + a.y * c);
    };
    t[1][1] = tw(t[1][1], c1, s1);
    t[1][2] = tw(t[1][2], c2, s2);
    t[2][1] = tw(t[2][1], c2, s2);
    t[2][2] = tw(t[2][2], c4, s4);
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
      double2 u[3] = {t[0][k1], t[1][k1], t[2][k1]};
      Dft<3, INV>::run(u);
      v[k1] = u[0];
      v[k1 + 3] = u[1];
      v[k1 + 6] = u[2];
    }
  }
};

// ---- one stage for the G threads of a line; Load(m) / Store(m, v) address the
//      line's elements.  SYNC: a barrier separates this stage's loads from its
//      stores (needed when both go to the same LDS buffer). ----
// butterflies per thread of stage S (at most 4: G = N / rmax, radix >= 2)
template <int N, int S, int G>
constexpr int stage_bpt() {
  return (N / make_plan(N).r[S] + G - 1) / G;
}

// The first twiddle w1 of each butterfly a thread runs in stage S, read from
// the table ahead of the transform (it does not depend on the data), so no
// stage waits on a global load.
template <int N, int S, int G, bool INV>
__device__ __forceinline__ void stage_twiddles(int g, const double2 *__restrict__ tw, double2 (&w)[4]) {
  constexpr PlanC P = make_plan(N);
  constexpr int R = P.r[S];
  constexpr int NS = plan_span(N, S);
  constexpr int NB = N / R;
  constexpr int BPT = stage_bpt<N, S, G>();
  static_assert(BPT <= 4, "at most four butterflies per thread");
#pragma unroll
  for (int i = 0; i < BPT; ++i) {
    const int b = g + i * G;
    const int k = (BPT * G == NB || b < NB) ? b % NS : 0;
    w[i] = NS > 1 ? tw[k * (N / (NS * R))] : make_double2(1.0, 0.0);
    if (INV) w[i].y = -w[i].y;
  }
}

// stage S's inputs of one thread: v[i][r] = element b + r * NB of butterfly
// b = g + i * G
template <int N, int S, int G>
using StageRegs = double2[stage_bpt<N, S, G>()][make_plan(N).r[S]];

template <int N, int S, int G, class Load>
__device__ __forceinline__ void stage_load(int g, const Load &load, StageRegs<N, S, G> &v) {
  constexpr int R = make_plan(N).r[S];
  constexpr int NB = N / R;
  constexpr int BPT = stage_bpt<N, S, G>();
#pragma unroll
  for (int i = 0; i < BPT; ++i) {
    const int b = g + i * G;
    if (BPT * G == NB || b < NB) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = load(b + r * NB);
    }
  }
}

template <int N, int S, int G, bool INV, class Store>
__device__ __forceinline__ void stage_exec(int g, const double2 (&wpre)[4], StageRegs<N, S, G> &v,
                                           const Store &store) {
  constexpr int R = make_plan(N).r[S];
  constexpr int NS = plan_span(N, S);
  constexpr int NB = N / R;
  constexpr int BPT = stage_bpt<N, S, G>();
#pragma unroll
  for (int i = 0; i < BPT; ++i) {
    const int b = g + i * G;
    if (BPT * G == NB || b < NB) {
      const int k = b % NS;
      if (NS > 1) {
        const double2 w1 = wpre[i];
        double2 w = w1;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          v[i][r] = cmul(v[i][r], w);
          if (r + 1 < R) w = cmul(w, w1);
        }
      }
      Dft<R, INV>::run(v[i]);
      const int idxD = (b / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) store(idxD + r * NS, v[i][r]);
    }
  }
}

template <int N, int S, int G, bool INV, bool SYNC, class Load, class Store>
__device__ __forceinline__ void stage(int g, const double2 (&wpre)[4], const Load &load, const Store &store) {
  StageRegs<N, S, G> v;
  stage_load<N, S, G>(g, load, v);
  if (SYNC) __syncthreads();
  stage_exec<N, S, G, INV>(g, wpre, v, store);
}

// Full transform of one line group.  Stage 0 loads through load0, the last
// stage stores through storeL; intermediate data lives in LDS through lds(m).
// LOAD0_LDS: load0 itself reads the LDS buffer, so stage 0 needs a barrier
// between its loads and its stores.
// the first twiddles of every stage a thread runs
template <int N, int G, bool INV>
__device__ __forceinline__ void load_twiddles(int g, const double2 *__restrict__ tw, double2 (&W)[7][4]) {
  constexpr PlanC P = make_plan(N);
  static_assert(P.ns <= 7, "too many stages");
  stage_twiddles<N, 0, G, INV>(g, tw, W[0]);
  if constexpr (P.ns > 1) stage_twiddles<N, (P.ns > 1 ? 1 : 0), G, INV>(g, tw, W[1]);
  if constexpr (P.ns > 2) stage_twiddles<N, (P.ns > 2 ? 2 : 0), G, INV>(g, tw, W[2]);
  if constexpr (P.ns > 3) stage_twiddles<N, (P.ns > 3 ? 3 : 0), G, INV>(g, tw, W[3]);
  if constexpr (P.ns > 4) stage_twiddles<N, (P.ns > 4 ? 4 : 0), G, INV>(g, tw, W[4]);
  if constexpr (P.ns > 5) stage_twiddles<N, (P.ns > 5 ? 5 : 0), G, INV>(g, tw, W[5]);
  if constexpr (P.ns > 6) stage_twiddles<N, (P.ns > 6 ? 6 : 0), G, INV>(g, tw, W[6]);
}

// Stages 1 .. ns-1 of a transform whose stage 0 has stored into LDS (and a
// barrier has passed); the last stage stores through storeL.
template <int N, int G, bool INV, class LdsAt, class StoreL>
__device__ __forceinline__ void run_rest(int g, const double2 (&W)[7][4], const LdsAt &lds, const StoreL &storeL) {
  constexpr PlanC P = make_plan(N);
  static_assert(P.ns >= 2, "a single-stage transform has no later stages");
  auto ldsload = [&](int m) { return lds(m); };
  auto ldsstore = [&](int m, double2 v) { lds(m) = v; };
  if constexpr (P.ns > 2) {
    stage<N, 1, G, INV, true>(g, W[1], ldsload, ldsstore);
    __syncthreads();
  }
  if constexpr (P.ns > 3) {
    stage<N, 2, G, INV, true>(g, W[2], ldsload, ldsstore);
    __syncthreads();
  }
  if constexpr (P.ns > 4) {
    stage<N, 3, G, INV, true>(g, W[3], ldsload, ldsstore);
    __syncthreads();
  }
  if constexpr (P.ns > 5) {
    stage<N, 4, G, INV, true>(g, W[4], ldsload, ldsstore);
    __syncthreads();
  }
  if constexpr (P.ns > 6) {
    stage<N, 5, G, INV, true>(g, W[5], ldsload, ldsstore);
    __syncthreads();
  }
  stage<N, P.ns - 1, G, INV, true>(g, W[P.ns - 1], ldsload, storeL);
}

// the whole transform with the first twiddles already in registers
template <int N, int G, bool INV, bool LOAD0_LDS = false, class Load0, class LdsAt, class StoreL>
__device__ __forceinline__ void run_fft_w(int g, const double2 (&W)[7][4], const Load0 &load0, const LdsAt &lds,
                                          const StoreL &storeL) {
  constexpr PlanC P = make_plan(N);
  if constexpr (P.ns == 1) {
    stage<N, 0, G, INV, false>(g, W[0], load0, storeL);
  } else {
    auto ldsstore = [&](int m, double2 v) { lds(m) = v; };
    stage<N, 0, G, INV, LOAD0_LDS>(g, W[0], load0, ldsstore);
    __syncthreads();
    run_rest<N, G, INV>(g, W, lds, storeL);
  }
}

template <int N, int G, bool INV, bool LOAD0_LDS = false, class Load0, class LdsAt, class StoreL>
__device__ __forceinline__ void run_fft(int g, const double2 *__restrict__ tw, const Load0 &load0, const LdsAt &lds,
                                        const StoreL &storeL) {
  double2 W[7][4];
  load_twiddles<N, G, INV>(g, tw, W);
  run_fft_w<N, G, INV, LOAD0_LDS>(g, W, load0, lds, storeL);
}

}  // namespace f2

// threads per line: the largest-radix stage has one butterfly per thread (a
// radix-9 stage counts as 8: its N/9 butterflies fit the N/8 threads of the
// radix-8 stages)
template <int N>
constexpr int f2_threads() {
  return N / (f2::plan_rmax(N) > 8 ? 8 : f2::plan_rmax(N));
}

// x-pass workgroup: 256 threads (several lines) or one line of G > 256 threads
template <int N>
constexpr int f2_xblock() {
  return f2_threads<N>() > 256 ? f2_threads<N>() : 256;
}

// LDS layout of an x-pass line: one padding value after every 8 (bank
// spreading), after every 16 for the 4608-point lines, so that such a line
// (78 KB) leaves room for a second workgroup on the CU's 160 KB
template <int N>
__device__ __forceinline__ int xpad(int m) {
  return N == 4608 ? m + (m >> 4) : f2::pad8(m);
}
template <int N>
constexpr int xpadded() {
  return N == 4608 ? N + N / 16 + 1 : f2::padded(N);
}

// The half spectrum of a real row of 2N points from the N-point transform Z
// of its packed pairs (in ls): X_k = (S_k - i w^k D_k) / 2 with
// S_k = Z_k + conj Z_{N-k}, D_k = Z_k - conj Z_{N-k}, w = exp(-2 pi i / 2N);
// values past k = N (pitch padding) are zero.
template <int N, int G, bool STREAM, typename S>
__device__ __forceinline__ void x_untangle_store(int g, const double2 *ls, const double2 *__restrict__ wx,
                                                 S *__restrict__ out, int pitch) {
  // pitch <= N + 32: a thread stores at most NK values.  Their twiddles are
  // all loaded before the first is used, so the thread waits once on them
  // (a load per iteration waited on every load in flight, the next line's
  // prefetch included, once per value).
  constexpr int NK = (N + 32 + G - 1) / G;
  double2 w[NK];
#pragma unroll
  for (int t = 0; t < NK; ++t) w[t] = wx[min(g + t * G, N)];
#pragma unroll
  for (int t = 0; t < NK; ++t) {
    const int kx = g + t * G;
    if (kx >= pitch) break;
    double2 X = make_double2(0.0, 0.0);
    if (kx <= N) {
      const double2 zk = ls[xpad<N>(kx == N ? 0 : kx)];
      double2 zc = ls[xpad<N>(kx == 0 ? 0 : N - kx)];
      zc.y = -zc.y;
      const double2 s = f2::cadd(zk, zc), d = f2::csub(zk, zc);
      const double2 wd = f2::cmul(w[t], d);
      X = make_double2(0.5 * (s.x + wd.y), 0.5 * (s.y - wd.x));
    }
    f2::st2x<STREAM>(out[kx], X);
  }
}

// The inverse: element m of the N-point inverse transform's input from the
// half spectrum X (in ls): X_m + conj X_{N-m} + i conj(w^m) (X_m - conj X_{N-m}).
template <int N>
__device__ __forceinline__ double2 x_retangle(const double2 *ls, const double2 *__restrict__ wx, int m) {
  const double2 xk = ls[xpad<N>(m)];
  double2 xc = ls[xpad<N>(N - m)];
  xc.y = -xc.y;
  const double2 s = f2::cadd(xk, xc), d = f2::csub(xk, xc);
  double2 w = wx[m];
  w.y = -w.y;
  const double2 wd = f2::cmul(w, d);
  return make_double2(s.x - wd.y, s.y + wd.x);
}

// Waves per SIMD the one-shot x kernels are held to.  2304-point lines: 288
// threads and 41 KB of LDS, three workgroups (15 waves) per CU at four waves
// per SIMD (128 VGPRs; left alone the inverse takes 129 and two fit).
// 4608-point lines: 576 threads and 78 KB, two workgroups (18 waves) per CU
// at five (102 VGPRs).
template <int N>
constexpr int f2_x_min_waves() {
  return N == 2304 ? 4 : (N == 4608 ? 5 : 1);
}
// The residual-forming form holds twenty more loads per point pair: held to
// those bounds it spills 96-188 bytes per lane, so it takes the registers it
// needs (three waves per SIMD at most).
template <int N, int SRC>
constexpr int f2_xfwd_min_waves() {
  return SRC == 2 ? ((N == 2304 || N == 4608) ? 3 : 1) : f2_x_min_waves<N>();
}

// ---- x forward: nlpb real rows of length 2N -> half spectra ---------------
// SRC = kXDiv: the rows are not read from src but formed from the velocities,
// as clcPrSrc does (src/m_TScheme.f90:48-77): sudtal times the divergence of
// (ux, uy, uz) with the 4th-order staggered x and z differences and the 2nd-
// order y difference, in k_prsrc's operand order.  Each workgroup then writes
// the maximum |divergence| over its rows to dv.part[blockIdx.x].
// SRC = kXRes: the rows are the refinement's FP64 residual f - L phi, in
// k_residual's operand order; each workgroup writes the sum of phi over its
// lines of row j = 1 to rs.psum[blockIdx.x].
constexpr int kXPlain = 0, kXDiv = 1, kXRes = 2;
template <int N, typename S, bool STREAM, int SRC>
__global__ void __launch_bounds__(f2_xblock<N>(), (f2_xfwd_min_waves<N, SRC>())) k2_fft_x_fwd(const double *__restrict__ src, S *__restrict__ spec,
                                                    const double2 *__restrict__ tw, const double2 *__restrict__ wx,
                                                    int ny, int nz, int j0, int nj, int pitch, long long rowlen,
                                                    DivSrc dv, ResSrc rs, Geo geo, int tj, int tk) {
  constexpr bool DIV = SRC == kXDiv, RES = SRC == kXRes, NBR = DIV || RES;
  constexpr int G = f2_threads<N>();
  constexpr int L = (256 / G) > 0 ? 256 / G : 1;
  constexpr int PN = xpadded<N>();
  __shared__ double2 sh[L * PN];
  const int l = threadIdx.x / G, g = threadIdx.x - l * G;
  // lines (j, k) of the planes j0 .. j0+nj-1 (0-based), all k.  Plain rows go
  // j fastest.  DIV reads the uz rows of planes k-1 .. k+2 and the uy rows j,
  // j+1 of each line, RES phi's rows of planes k-3 .. k+3 and j-1, j+1: their
  // lines go in tiles of tj x tk lines (tj | nj, tk | nz; k fastest inside a
  // tile, the tiles k fastest) and runs of two tiles' workgroups go to one XCD
  // (xcd_run), so the neighbouring lines that re-read a row, in j or in k, do
  // so through the same L2.
  const long long nlines = (long long)nj * nz;
  const int tt = tj * tk;
  const long long wb = NBR ? (long long)xcd_run((int)blockIdx.x, (int)gridDim.x, tt >= L ? 2 * tt / L : 1)
                           : (long long)blockIdx.x;
  const long long q = wb * L + l;
  const bool live = l < L && q < nlines;
  const long long qq = live ? q : 0;
  const long long tq = qq / tt, wq = qq - tq * tt, ntk = nz / tk;
  const long long jl = NBR ? j0 + (tq / ntk) * tj + wq / tk : j0 + qq % nj;
  const long long k = NBR ? (tq % ntk) * tk + wq % tk : qq / nj;
  const long long sq = jl + (long long)ny * k;  // spectral row S[k][jl]
  const double2 *row = reinterpret_cast<const double2 *>(src + rowlen * (jl + 1 + (long long)(ny + 2) * k));
  double2 *ls = sh + (l < L ? l : 0) * PN;
  double dmax = 0.0;
  auto load_div = [&](int m) {
    // the pair of points (2m, 2m+1) of row j = jl + 1, plane k
    const long long rj = rowlen * (jl + 1 + (long long)(ny + 2) * k);
    const double2 *uxr = reinterpret_cast<const double2 *>(dv.ux + rj);
    const double2 *uyr = reinterpret_cast<const double2 *>(dv.uy + rj);
    const int mm = m > 0 ? m - 1 : N - 1, mp = m + 1 < N ? m + 1 : 0;
    const double2 a = uxr[mm], b = uxr[m], cc = uxr[mp];
    const double2 y0 = uyr[m], y1 = uyr[m + N];
    const long long pl = rowlen * (ny + 2);
    double2 z[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      long long kk = k - 1 + s;
      kk = kk < 0 ? kk + nz : (kk >= nz ? kk - nz : kk);
      z[s] = reinterpret_cast<const double2 *>(dv.uz + rj + (kk - k) * pl)[m];
    }
    const double rdyp = geo.met[M_RDYP * geo.nyh + jl + 1];
    const double dudx0 = geo.dxc1 * a.y + geo.dxc2 * b.x + geo.dxc3 * b.y + geo.dxc4 * cc.x;
    const double dudx1 = geo.dxc1 * b.x + geo.dxc2 * b.y + geo.dxc3 * cc.x + geo.dxc4 * cc.y;
    const double dvdy0 = (y1.x - y0.x) * rdyp, dvdy1 = (y1.y - y0.y) * rdyp;
    const double dwdz0 = geo.dzc1 * z[0].x + geo.dzc2 * z[1].x + geo.dzc3 * z[2].x + geo.dzc4 * z[3].x;
    const double dwdz1 = geo.dzc1 * z[0].y + geo.dzc2 * z[1].y + geo.dzc3 * z[2].y + geo.dzc4 * z[3].y;
    const double r0 = dudx0 + dvdy0 + dwdz0, r1 = dudx1 + dvdy1 + dwdz1;
    dmax = fmax(dmax, fmax(fabs(r0), fabs(r1)));
    return make_double2(dv.sudtal * r0, dv.sudtal * r1);
  };
  double psum = 0.0;
  auto load_res = [&](int m) {
    // the pair of points (2m, 2m+1) of row j = jl + 1, plane k
    constexpr double st0 = -1460.0 / 576.0, st1 = 783.0 / 576.0, st2 = -54.0 / 576.0, st3 = 1.0 / 576.0;
    const long long rj = rowlen * (jl + 1 + (long long)(ny + 2) * k);
    const double2 *pr = reinterpret_cast<const double2 *>(rs.phi + rj);
    double2 p[5];  // pairs m-2 .. m+2: points 2m-4 .. 2m+5
#pragma unroll
    for (int s = 0; s < 5; ++s) {
      int mm = m + s - 2;
      mm = mm < 0 ? mm + N : (mm >= N ? mm - N : mm);
      p[s] = pr[mm];
    }
    const long long pl = rowlen * (ny + 2);
    double2 z[7];
#pragma unroll
    for (int s = 0; s < 7; ++s) {
      long long kk = k - 3 + s;
      kk = kk < 0 ? kk + nz : (kk >= nz ? kk - nz : kk);
      z[s] = reinterpret_cast<const double2 *>(rs.phi + rj + (kk - k) * pl)[m];
    }
    const int j = (int)jl + 1;
    const double2 c0 = p[2];
    const double2 ym = j > 1 ? pr[m - N] : c0, yp = j < ny ? pr[m + N] : c0;
    const double2 fv = reinterpret_cast<const double2 *>(rs.f + rj)[m];
    const double am = geo.met[M_AM2PH * geo.nyh + j], ac = geo.met[M_AC2PH * geo.nyh + j],
                 ap = geo.met[M_AP2PH * geo.nyh + j];
    const double lx0 = st0 * c0.x + st1 * (p[1].y + c0.y) + st2 * (p[1].x + p[3].x) + st3 * (p[0].y + p[3].y);
    const double lx1 = st0 * c0.y + st1 * (c0.x + p[3].x) + st2 * (p[1].y + p[3].y) + st3 * (p[1].x + p[4].x);
    const double lz0 = st0 * c0.x + st1 * (z[2].x + z[4].x) + st2 * (z[1].x + z[5].x) + st3 * (z[0].x + z[6].x);
    const double lz1 = st0 * c0.y + st1 * (z[2].y + z[4].y) + st2 * (z[1].y + z[5].y) + st3 * (z[0].y + z[6].y);
    const double lap0 = rs.rdx2 * lx0 + rs.rdz2 * lz0 + (am * ym.x + ac * c0.x + ap * yp.x);
    const double lap1 = rs.rdx2 * lx1 + rs.rdz2 * lz1 + (am * ym.y + ac * c0.y + ap * yp.y);
    if (j == 1) {
      psum += c0.x;
      psum += c0.y;
    }
    return make_double2(fv.x - lap0, fv.y - lap1);
  };
  auto load0 = [&](int m) {
    if (!live) return make_double2(0.0, 0.0);
    if constexpr (DIV) return load_div(m);
    else if constexpr (RES) return load_res(m);
    else return f2::ld2x<STREAM>(row[m]);
  };
  auto lds = [&](int m) -> double2 & { return ls[xpad<N>(m)]; };
  auto store_lds = [&](int m, double2 v) { ls[xpad<N>(m)] = v; };
  f2::run_fft<N, G, false>(g, tw, load0, lds, store_lds);
  __syncthreads();
  if (live) x_untangle_store<N, G, STREAM>(g, ls, wx, spec + sq * pitch, pitch);
  if constexpr (DIV) {
    // fixed-order tree over the workgroup, in the FFT's LDS once its reads are done
    double *red = reinterpret_cast<double *>(sh);
    __syncthreads();
    red[threadIdx.x] = dmax;
    __syncthreads();
    for (int n = blockDim.x; n > 1;) {
      const int h = (n + 1) / 2;
      if ((int)threadIdx.x < n - h) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + h]);
      __syncthreads();
      n = h;
    }
    if (threadIdx.x == 0) dv.part[blockIdx.x] = red[0];
  }
  if constexpr (RES) {
    // fixed-order tree over the workgroup, in the FFT's LDS once its reads are done
    double *red = reinterpret_cast<double *>(sh);
    __syncthreads();
    red[threadIdx.x] = psum;
    __syncthreads();
    for (int n = blockDim.x; n > 1;) {
      const int h = (n + 1) / 2;
      if ((int)threadIdx.x < n - h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
      n = h;
    }
    if (threadIdx.x == 0) rs.psum[blockIdx.x] = red[0];
  }
}

// ---- x inverse: half spectra -> real rows times scale ----------------------
// ACC: add the result to dst (a correction) instead of storing it, minus the
// device scalar *shift (the mean the correction removes).  ghost: the rows
// j = 1 and j = ny also go to the Neumann ghost rows j = 0 and j = ny+1
// (SetBC_and_UpdateHalo_pr, src/m_BC_and_Halo.f90:115-120).

template <int N, typename S, bool ACC, bool STREAM>
__global__ void __launch_bounds__(f2_xblock<N>(), f2_x_min_waves<N>())
    k2_fft_x_inv(const S *__restrict__ spec, double *__restrict__ dst,
                                                    const double2 *__restrict__ tw, const double2 *__restrict__ wx,
                                                    int ny, int nz, int j0, int nj, int pitch, long long rowlen,
                                                    double scale, const double *__restrict__ shift, int ghost) {
  constexpr int G = f2_threads<N>();
  constexpr int L = (256 / G) > 0 ? 256 / G : 1;
  constexpr int PN = xpadded<N>();
  __shared__ double2 sh[L * PN];
  const int l = threadIdx.x / G, g = threadIdx.x - l * G;
  const long long nlines = (long long)nj * nz;
  const long long q = (long long)blockIdx.x * L + l;
  const bool live = l < L && q < nlines;
  const long long qq = live ? q : 0;
  const long long jl = j0 + qq % nj, k = qq / nj;
  const S *in = spec + (jl + (long long)ny * k) * pitch;
  double2 *ls = sh + (l < L ? l : 0) * PN;
  // The re-tangle of element m needs X_m and X_{N-m}: the line's N+1 values
  // go through LDS once (coalesced) instead of being fetched twice.  All of a
  // thread's loads are issued before the first LDS write (lines past the end
  // read line 0 and store nothing; indices past N repeat element N, so
  // every thread that reaches it writes the same value).
  constexpr int NT = N / G + 1;
  double2 pre[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) pre[t] = f2::ld2x<STREAM>(in[min(g + t * G, N)]);
#pragma unroll
  for (int t = 0; t < NT; ++t) ls[xpad<N>(min(g + t * G, N))] = pre[t];
  __syncthreads();
  auto load0 = [&](int m) { return x_retangle<N>(ls, wx, m); };
  auto lds = [&](int m) -> double2 & { return ls[xpad<N>(m)]; };
  double2 *row = reinterpret_cast<double2 *>(dst + rowlen * (jl + 1 + (long long)(ny + 2) * k));
  // the ghost row this line also fills (or none)
  double2 *grow = nullptr;
  if (ghost && jl == 0) grow = reinterpret_cast<double2 *>(dst + rowlen * ((long long)(ny + 2) * k));
  if (ghost && jl == ny - 1) grow = reinterpret_cast<double2 *>(dst + rowlen * (ny + 1 + (long long)(ny + 2) * k));
  const double sh0 = ACC ? shift[0] : 0.0;
  auto store_out = [&](int m, double2 v) {
    if (!live) return;
    double2 x;
    if (ACC) {
      const double2 o = f2::ld2x<STREAM>(row[m]);
      x = make_double2(o.x + v.x * scale - sh0, o.y + v.y * scale - sh0);
    } else {
      x = make_double2(v.x * scale, v.y * scale);
    }
    f2::st2x<STREAM>(row[m], x);
    if (grow) f2::st2x<STREAM>(grow[m], x);
  };
  f2::run_fft<N, G, true, true>(g, tw, load0, lds, store_out);
}

// ---- z pass: T columns x N rows, thread id = g * T + col ----------------
// 1-D grid of ntile * nj workgroups: column tile and plane from the
// XCD-grouped index, so that the tiles sharing a 128-byte row segment (narrow
// tiles of the long lengths) are read and written through one L2.
template <int N, int T, bool INV, typename S>
__global__ void __launch_bounds__(f2_threads<N>() * T) k2_fft_z(S *__restrict__ spec, const double2 *__restrict__ tw, int ny,
                                                 int pitch, int j0, int ntile) {
  constexpr int G = f2_threads<N>();
  constexpr int GRP = (int)(128 / (T * sizeof(S))) > 1 ? (int)(128 / (T * sizeof(S))) : 1;
  __shared__ double2 sh[N * T];
  const int col = threadIdx.x % T, g = threadIdx.x / T;
  const long long zs = (long long)pitch * ny;
  const int w = xcd_group<GRP>((int)blockIdx.x, (int)gridDim.x);
  const int tile = w % ntile, jj = w / ntile;
  S *base = spec + (long long)(j0 + jj) * pitch + (long long)tile * T + col;
  auto load0 = [&](int m) { return f2::ld2(base[m * zs]); };
  auto lds = [&](int m) -> double2 & { return sh[m * T + col]; };
  auto store_out = [&](int m, double2 v) { f2::st2(base[m * zs], v); };
  f2::run_fft<N, G, INV>(g, tw, load0, lds, store_out);
}

// ---- persistent passes with a register prefetch of the next work item ------
// The long lengths keep one workgroup per CU (their lines fill 83-90 KB of
// LDS), so a one-shot workgroup idles through the HBM latency of its own
// loads and the CU has no other workgroup to switch to.  These kernels run
// one or a few workgroups per CU over the work items (line groups, or z
// column tiles) in strides of the grid; each issues the global loads of its
// next item -- the inputs of stage 0, straight into the registers stage 0
// of the current item has just consumed -- before the remaining stages, so
// those loads are in flight during the LDS stages.  The barriers wait on LDS traffic only (s_waitcnt lgkmcnt), not
// on the outstanding global loads.  Same arithmetic as the one-shot kernels.

template <int N, typename S, bool STREAM>
__global__ void __launch_bounds__(f2_xblock<N>()) k2p_fft_x_fwd(const double *__restrict__ src, S *__restrict__ spec,
                                                     const double2 *__restrict__ tw, const double2 *__restrict__ wx,
                                                     int ny, int nz, int j0, int nj, int pitch, long long rowlen) {
  constexpr int G = f2_threads<N>();
  constexpr int L = (256 / G) > 0 ? 256 / G : 1;
  constexpr int PN = xpadded<N>();
  constexpr int R0 = f2::make_plan(N).r[0];
  constexpr int NB0 = N / R0;
  constexpr int BPT0 = f2::stage_bpt<N, 0, G>();
  __shared__ double2 sh[L * PN];
  const int l = threadIdx.x / G, g = threadIdx.x - l * G;
  const long long nlines = (long long)nj * nz;
  const long long ngroups = (nlines + L - 1) / L;
  double2 *ls = sh + l * PN;
  // line l of group grp: rows j fastest; sq = its spectral row
  auto line = [&](long long grp, bool &live, long long &sq) {
    const long long q = grp * L + l;
    live = q < nlines;
    const long long qq = live ? q : 0;
    const long long jl = j0 + qq % nj, k = qq / nj;
    sq = jl + (long long)ny * k;
    return reinterpret_cast<const double2 *>(src + rowlen * (jl + 1 + (long long)(ny + 2) * k));
  };
  auto fetch = [&](long long grp, f2::StageRegs<N, 0, G> &v) {
    bool live;
    long long sq;
    const double2 *row = line(grp, live, sq);
#pragma unroll
    for (int i = 0; i < BPT0; ++i) {
      const int b = g + i * G;
      if (BPT0 * G == NB0 || b < NB0) {
#pragma unroll
        for (int r = 0; r < R0; ++r) v[i][r] = live ? f2::ld2x<STREAM>(row[b + r * NB0]) : make_double2(0.0, 0.0);
      }
    }
  };
  auto lds = [&](int m) -> double2 & { return ls[xpad<N>(m)]; };
  auto store_lds = [&](int m, double2 v) { ls[xpad<N>(m)] = v; };
  f2::StageRegs<N, 0, G> v;
  long long grp = blockIdx.x;
  if (grp >= ngroups) return;
  fetch(grp, v);
  for (; grp < ngroups; grp += gridDim.x) {
    // the lane index through an empty asm: the per-stage LDS addresses are
    // recomputed in each iteration, as in the one-shot kernel, rather than
    // hoisted out of the loop into registers
    int gi = g;
    asm volatile("" : "+v"(gi));
    // the stage twiddles, reloaded per item (cache hits) rather than held
    // across the loop
    double2 W[7][4];
    f2::load_twiddles<N, G, false>(gi, tw, W);
    // stage 0 consumes the registers; the next group's loads then go into
    // them (the last group re-reads its own: no branch around the loads, and
    // the values are discarded).  The scheduling barrier keeps those loads
    // after stage 0, so the two sets of registers are never live together.
    f2::stage_exec<N, 0, G, false>(gi, W[0], v, store_lds);
    __builtin_amdgcn_sched_barrier(0);
    const long long gn = grp + gridDim.x < ngroups ? grp + gridDim.x : grp;
    fetch(gn, v);
    __syncthreads();
    f2::run_rest<N, G, false>(gi, W, lds, store_lds);
    __syncthreads();
    bool live;
    long long sq;
    (void)line(grp, live, sq);
    if (live) x_untangle_store<N, G, STREAM>(gi, ls, wx, spec + sq * pitch, pitch);
    __syncthreads();
  }
}

template <int N, typename S, bool ACC, bool STREAM>
__global__ void __launch_bounds__(f2_xblock<N>()) k2p_fft_x_inv(const S *__restrict__ spec, double *__restrict__ dst,
                                                     const double2 *__restrict__ tw, const double2 *__restrict__ wx,
                                                     int ny, int nz, int j0, int nj, int pitch, long long rowlen,
                                                     double scale, const double *__restrict__ shift, int ghost) {
  constexpr int G = f2_threads<N>();
  constexpr int L = (256 / G) > 0 ? 256 / G : 1;
  constexpr int PN = xpadded<N>();
  constexpr int NT = N / G + 1;
  __shared__ double2 sh[L * PN];
  const int l = threadIdx.x / G, g = threadIdx.x - l * G;
  const long long nlines = (long long)nj * nz;
  const long long ngroups = (nlines + L - 1) / L;
  double2 *ls = sh + l * PN;
  const double sh0 = ACC ? shift[0] : 0.0;
  auto coords = [&](long long grp, bool &live, long long &jl, long long &k) {
    const long long q = grp * L + l;
    live = q < nlines;
    const long long qq = live ? q : 0;
    jl = j0 + qq % nj;
    k = qq / nj;
  };
  // the line's N+1 spectral values (indices past N repeat element N)
  auto fetch = [&](long long grp, double2 (&v)[NT]) {
    bool live;
    long long jl, k;
    coords(grp, live, jl, k);
    const S *in = spec + (jl + (long long)ny * k) * pitch;
#pragma unroll
    for (int t = 0; t < NT; ++t) v[t] = f2::ld2x<STREAM>(in[min(g + t * G, N)]);
  };
  auto lds = [&](int m) -> double2 & { return ls[xpad<N>(m)]; };
  double2 v[NT];
  long long grp = blockIdx.x;
  if (grp >= ngroups) return;
  fetch(grp, v);
  for (; grp < ngroups; grp += gridDim.x) {
    int gi = g;  // recomputed addresses and twiddles per iteration, as in the forward pass
    asm volatile("" : "+v"(gi));
    double2 W[7][4];
    f2::load_twiddles<N, G, true>(gi, tw, W);
#pragma unroll
    for (int t = 0; t < NT; ++t) ls[xpad<N>(min(gi + t * G, N))] = v[t];
    // the registers are free: the next group's loads go into them
    __builtin_amdgcn_sched_barrier(0);
    const long long gn = grp + gridDim.x < ngroups ? grp + gridDim.x : grp;
    fetch(gn, v);
    __syncthreads();
    bool live;
    long long jl, k;
    coords(grp, live, jl, k);
    double2 *row = reinterpret_cast<double2 *>(dst + rowlen * (jl + 1 + (long long)(ny + 2) * k));
    double2 *grow = nullptr;
    if (ghost && jl == 0) grow = reinterpret_cast<double2 *>(dst + rowlen * ((long long)(ny + 2) * k));
    if (ghost && jl == ny - 1) grow = reinterpret_cast<double2 *>(dst + rowlen * (ny + 1 + (long long)(ny + 2) * k));
    auto load0 = [&](int m) { return x_retangle<N>(ls, wx, m); };
    auto store_out = [&](int m, double2 v) {
      if (!live) return;
      double2 x;
      if (ACC) {
        const double2 o = f2::ld2x<STREAM>(row[m]);
        x = make_double2(o.x + v.x * scale - sh0, o.y + v.y * scale - sh0);
      } else {
        x = make_double2(v.x * scale, v.y * scale);
      }
      f2::st2x<STREAM>(row[m], x);
      if (grow) f2::st2x<STREAM>(grow[m], x);
    };
    f2::run_fft_w<N, G, true, true>(gi, W, load0, lds, store_out);
    __syncthreads();
  }
}

// z pass over nitems = ntile * nj column tiles (XCD-grouped as the one-shot
// kernel, so the tiles that share a 128-byte row segment are in flight on
// one XCD together)
template <int N, int T, bool INV, typename S>
__global__ void __launch_bounds__(f2_threads<N>() * T) k2p_fft_z(S *__restrict__ spec, const double2 *__restrict__ tw,
                                                  int ny, int pitch, int j0, int ntile, int nitems) {
  constexpr int G = f2_threads<N>();
  constexpr int GRP = (int)(128 / (T * sizeof(S))) > 1 ? (int)(128 / (T * sizeof(S))) : 1;
  constexpr int R0 = f2::make_plan(N).r[0];
  constexpr int NB0 = N / R0;
  constexpr int BPT0 = f2::stage_bpt<N, 0, G>();
  __shared__ double2 sh[N * T];
  const int col = threadIdx.x % T, g = threadIdx.x / T;
  const long long zs = (long long)pitch * ny;
  auto base_of = [&](int item) {
    const int w = xcd_group<GRP>(item, nitems);
    const int tile = w % ntile, jj = w / ntile;
    return spec + (long long)(j0 + jj) * pitch + (long long)tile * T + col;
  };
  auto fetch = [&](int item, f2::StageRegs<N, 0, G> &v) {
    const S *base = base_of(item);
#pragma unroll
    for (int i = 0; i < BPT0; ++i) {
      const int b = g + i * G;
      if (BPT0 * G == NB0 || b < NB0) {
#pragma unroll
        for (int r = 0; r < R0; ++r) v[i][r] = f2::ld2(base[(long long)(b + r * NB0) * zs]);
      }
    }
  };
  f2::StageRegs<N, 0, G> v;
  int item = blockIdx.x;
  if (item >= nitems) return;
  fetch(item, v);
  for (; item < nitems; item += gridDim.x) {
    // lane indices through an empty asm: addresses recomputed per iteration
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int ci = tid % T, gi = tid / T;
    double2 W[7][4];
    f2::load_twiddles<N, G, INV>(gi, tw, W);
    auto lds = [&](int m) -> double2 & { return sh[m * T + ci]; };
    auto store_lds = [&](int m, double2 x) { sh[m * T + ci] = x; };
    f2::stage_exec<N, 0, G, INV>(gi, W[0], v, store_lds);
    __builtin_amdgcn_sched_barrier(0);
    const int in = item + (int)gridDim.x < nitems ? item + (int)gridDim.x : item;
    fetch(in, v);
    S *base = base_of(item);
    auto store_out = [&](int m, double2 x) { f2::st2(base[(long long)m * zs], x); };
    __syncthreads();
    f2::run_rest<N, G, INV>(gi, W, lds, store_out);
    __syncthreads();
  }
}

// ---- dispatch ----------------------------------------------------------------
// lengths with persistent forms: the x lengths of the 1024-, 4608- and
// 9216-point rows, the z length 1400 with its 2- and 4-column tiles.  The
// shorter z lengths run faster one-shot: at 1024x256x512 the 8-column pass
// took 0.416 ms against 0.438 (16 columns, 44 bytes of scratch per lane at the
// 128 VGPRs of a 1024-thread workgroup, 1.34x the read bytes) and 0.456 (8
// columns) persistent; at 4608x140x700 1.77 ms against 1.80-1.82 (8 columns,
// 28 bytes of scratch per lane) and 2.18 (4 columns).
constexpr bool has_px(int n) { return n == 512 || n == 2304 || n == 4608; }
constexpr bool has_pz(int n, int t) { return n == 1400 && (t == 2 || t == 4); }

// workgroups of a persistent kernel resident per CU (at least one)
template <class K>
static int resident_blocks(K kernel, int threads) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, 0) != hipSuccess || nb < 1) {
    (void)hipGetLastError();
    nb = 1;
  }
  return nb;
}

static unsigned persistent_grid(const ptc_ctx *c, int per_cu, long long items) {
  return (unsigned)std::max<long long>(1, std::min<long long>(items, (long long)per_cu * std::max(1, c->n_cu)));
}

template <int N, typename S, bool STREAM>
static void px_fwd_launch(ptc_ctx *c, const double *src, S *spec, int j0, int nj) {
  constexpr int G = f2_threads<N>();
  constexpr int L = (256 / G) > 0 ? 256 / G : 1;
  static const int per_cu = resident_blocks(k2p_fft_x_fwd<N, S, STREAM>, L * G);
  const long long groups = ((long long)nj * c->nz + L - 1) / L;
  hipLaunchKernelGGL((k2p_fft_x_fwd<N, S, STREAM>), dim3(persistent_grid(c, per_cu, groups)), dim3(L * G), 0,
                     c->stream, src, spec, c->pp.fx.tw, c->pp.d_wx, c->ny, c->nz, j0, nj, c->pp.pitch,
                     (long long)c->nx);
}

template <int N, typename S, bool ACC, bool STREAM>
static void px_inv_launch(ptc_ctx *c, const S *spec, double *dst, double scale, int j0, int nj, bool ghost) {
  constexpr int G = f2_threads<N>();
  constexpr int L = (256 / G) > 0 ? 256 / G : 1;
  static const int per_cu = resident_blocks(k2p_fft_x_inv<N, S, ACC, STREAM>, L * G);
  const long long groups = ((long long)nj * c->nz + L - 1) / L;
  hipLaunchKernelGGL((k2p_fft_x_inv<N, S, ACC, STREAM>), dim3(persistent_grid(c, per_cu, groups)), dim3(L * G), 0,
                     c->stream, spec, dst, c->pp.fx.tw, c->pp.d_wx, c->ny, c->nz, j0, nj, c->pp.pitch,
                     (long long)c->nx, scale, c->d_scal + kSlotPinShift, ghost ? 1 : 0);
}

template <int N, int T, bool INV, typename S>
static void pz_launch(ptc_ctx *c, S *spec, int j0, int nj) {
  constexpr int G = f2_threads<N>();
  static const int per_cu = resident_blocks(k2p_fft_z<N, T, INV, S>, G * T);
  const int ntile = c->pp.pitch / T;
  const int nitems = ntile * nj;
  hipLaunchKernelGGL((k2p_fft_z<N, T, INV, S>), dim3(persistent_grid(c, per_cu, nitems)), dim3(G * T), 0, c->stream,
                     spec, c->pp.fz.tw, c->ny, c->pp.pitch, j0, ntile, nitems);
}

template <int N>
static long long x_fwd_blocks(long long nlines) {
  constexpr int G = f2_threads<N>();
  constexpr int L = (256 / G) > 0 ? 256 / G : 1;
  return (nlines + L - 1) / L;
}

template <int N, typename S, bool STREAM, int SRC>
static void x_fwd_launch(ptc_ctx *c, const double *src, S *spec, int j0, int nj, const DivSrc &dv,
                         const ResSrc &rs = ResSrc{}) {
  constexpr int G = f2_threads<N>();
  constexpr int L = (256 / G) > 0 ? 256 / G : 1;
  const unsigned nblk = (unsigned)x_fwd_blocks<N>((long long)nj * c->nz);
  // line tiles of the neighbour-reading forms: the largest of 8, 4, 2, 1 rows
  // dividing nj, of 16, 8, 4, 2, 1 planes dividing nz
  int tj = 8, tk = 16;
  while (nj % tj) tj /= 2;
  while (c->nz % tk) tk /= 2;
  hipLaunchKernelGGL((k2_fft_x_fwd<N, S, STREAM, SRC>), dim3(nblk), dim3(L * G), 0, c->stream, src, spec,
                     c->pp.fx.tw, c->pp.d_wx, c->ny, c->nz, j0, nj, c->pp.pitch, (long long)c->nx, dv, rs, c->geo,
                     tj, tk);
}

template <int N>
static int x_fwd_t(ptc_ctx *c, const double *src, void *spec, int j0, int nj, bool f64, const DivSrc *div,
                   const ResSrc *res) {
  const bool f32 = !f64 && c->pp.prec == PTC_PREC_FP32_REFINE;
  const DivSrc none{};
  if (div) {
    PTC_CHECK(!f32, PTC_E_STATE, "the divergence-forming x pass needs FP64 spectra");
    if (c->pp.stream) x_fwd_launch<N, double2, true, kXDiv>(c, src, (double2 *)spec, j0, nj, *div);
    else x_fwd_launch<N, double2, false, kXDiv>(c, src, (double2 *)spec, j0, nj, *div);
  } else if (res) {
    // the one-shot kernel forms the residual, whatever the plan's x form
    if (f32 && c->pp.stream) x_fwd_launch<N, float2, true, kXRes>(c, src, (float2 *)spec, j0, nj, none, *res);
    else if (f32) x_fwd_launch<N, float2, false, kXRes>(c, src, (float2 *)spec, j0, nj, none, *res);
    else if (c->pp.stream) x_fwd_launch<N, double2, true, kXRes>(c, src, (double2 *)spec, j0, nj, none, *res);
    else x_fwd_launch<N, double2, false, kXRes>(c, src, (double2 *)spec, j0, nj, none, *res);
  } else if (c->pp.persist_x) {
    if constexpr (has_px(N)) {
      if (f32 && c->pp.stream) px_fwd_launch<N, float2, true>(c, src, (float2 *)spec, j0, nj);
      else if (f32) px_fwd_launch<N, float2, false>(c, src, (float2 *)spec, j0, nj);
      else if (c->pp.stream) px_fwd_launch<N, double2, true>(c, src, (double2 *)spec, j0, nj);
      else px_fwd_launch<N, double2, false>(c, src, (double2 *)spec, j0, nj);
    } else {
      set_error("no persistent x pass for this length");
      return PTC_E_UNSUPPORTED;
    }
  } else if (f32 && c->pp.stream) {
    x_fwd_launch<N, float2, true, kXPlain>(c, src, (float2 *)spec, j0, nj, none);
  } else if (f32) {
    x_fwd_launch<N, float2, false, kXPlain>(c, src, (float2 *)spec, j0, nj, none);
  } else if (c->pp.stream) {
    x_fwd_launch<N, double2, true, kXPlain>(c, src, (double2 *)spec, j0, nj, none);
  } else {
    x_fwd_launch<N, double2, false, kXPlain>(c, src, (double2 *)spec, j0, nj, none);
  }
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

template <int N, typename S, bool ACC, bool STREAM>
static void x_inv_launch(ptc_ctx *c, const S *spec, double *dst, double scale, int j0, int nj, bool ghost) {
  constexpr int G = f2_threads<N>();
  constexpr int L = (256 / G) > 0 ? 256 / G : 1;
  const long long nlines = (long long)nj * c->nz;
  const unsigned nblk = (unsigned)((nlines + L - 1) / L);
  hipLaunchKernelGGL((k2_fft_x_inv<N, S, ACC, STREAM>), dim3(nblk), dim3(L * G), 0, c->stream, spec, dst,
                     c->pp.fx.tw, c->pp.d_wx, c->ny, c->nz, j0, nj, c->pp.pitch, (long long)c->nx, scale,
                     c->d_scal + kSlotPinShift, ghost ? 1 : 0);
}

template <int N>
static int x_inv_t(ptc_ctx *c, const void *spec, double *dst, double scale, int j0, int nj, bool acc, bool ghost) {
  const bool f32 = c->pp.prec == PTC_PREC_FP32_REFINE, st = c->pp.stream;
  const float2 *s4 = (const float2 *)spec;
  const double2 *s8 = (const double2 *)spec;
  if (c->pp.persist_x) {
    if constexpr (has_px(N)) {
      if (f32) {
        if (acc && st) px_inv_launch<N, float2, true, true>(c, s4, dst, scale, j0, nj, ghost);
        else if (acc) px_inv_launch<N, float2, true, false>(c, s4, dst, scale, j0, nj, ghost);
        else if (st) px_inv_launch<N, float2, false, true>(c, s4, dst, scale, j0, nj, ghost);
        else px_inv_launch<N, float2, false, false>(c, s4, dst, scale, j0, nj, ghost);
      } else {
        PTC_CHECK(!acc, PTC_E_STATE, "accumulating x-inverse with FP64 spectra");
        if (st) px_inv_launch<N, double2, false, true>(c, s8, dst, scale, j0, nj, ghost);
        else px_inv_launch<N, double2, false, false>(c, s8, dst, scale, j0, nj, ghost);
      }
      PTC_HIP(hipGetLastError());
      return PTC_OK;
    } else {
      set_error("no persistent x pass for this length");
      return PTC_E_UNSUPPORTED;
    }
  }
  if (f32) {
    if (acc && st) x_inv_launch<N, float2, true, true>(c, s4, dst, scale, j0, nj, ghost);
    else if (acc) x_inv_launch<N, float2, true, false>(c, s4, dst, scale, j0, nj, ghost);
    else if (st) x_inv_launch<N, float2, false, true>(c, s4, dst, scale, j0, nj, ghost);
    else x_inv_launch<N, float2, false, false>(c, s4, dst, scale, j0, nj, ghost);
  } else {
    // the accumulating form belongs to the FP32-storage refinement only
    PTC_CHECK(!acc, PTC_E_STATE, "accumulating x-inverse with FP64 spectra");
    if (st) x_inv_launch<N, double2, false, true>(c, s8, dst, scale, j0, nj, ghost);
    else x_inv_launch<N, double2, false, false>(c, s8, dst, scale, j0, nj, ghost);
  }
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

template <int N, int T>
static int z_t(ptc_ctx *c, void *spec, bool inverse, int j0, int nj, bool f64) {
  constexpr int G = f2_threads<N>();
  const int ntile = c->pp.pitch / T;
  dim3 grid((unsigned)(ntile * nj));
  const auto *tw = c->pp.fz.tw;
  const int ny = c->ny, pitch = c->pp.pitch;
  if (c->pp.persist_z) {
    if constexpr (has_pz(N, T)) {
      if (!f64 && c->pp.prec == PTC_PREC_FP32_REFINE) {
        if (inverse) pz_launch<N, T, true, float2>(c, (float2 *)spec, j0, nj);
        else pz_launch<N, T, false, float2>(c, (float2 *)spec, j0, nj);
      } else {
        if (inverse) pz_launch<N, T, true, double2>(c, (double2 *)spec, j0, nj);
        else pz_launch<N, T, false, double2>(c, (double2 *)spec, j0, nj);
      }
      PTC_HIP(hipGetLastError());
      return PTC_OK;
    } else {
      set_error("no persistent z pass for this length and tile");
      return PTC_E_UNSUPPORTED;
    }
  }
  if (!f64 && c->pp.prec == PTC_PREC_FP32_REFINE) {
    if (inverse)
      hipLaunchKernelGGL((k2_fft_z<N, T, true, float2>), grid, dim3(G * T), 0, c->stream, (float2 *)spec, tw, ny, pitch,
                         j0, ntile);
    else
      hipLaunchKernelGGL((k2_fft_z<N, T, false, float2>), grid, dim3(G * T), 0, c->stream, (float2 *)spec, tw, ny,
                         pitch, j0, ntile);
  } else {
    if (inverse)
      hipLaunchKernelGGL((k2_fft_z<N, T, true, double2>), grid, dim3(G * T), 0, c->stream, (double2 *)spec, tw, ny,
                         pitch, j0, ntile);
    else
      hipLaunchKernelGGL((k2_fft_z<N, T, false, double2>), grid, dim3(G * T), 0, c->stream, (double2 *)spec, tw, ny,
                         pitch, j0, ntile);
  }
  PTC_HIP(hipGetLastError());
  return PTC_OK;
}

// x lengths h = nx/2 with a specialised kernel
// (2304 and 4608: the 4608- and 9216-point rows of the largest cases; their
// lines take 41 KB and 83 KB of LDS, which gfx950's 160 KB allows)
#define PTC_X_SIZES(X) X(8) X(16) X(32) X(48) X(64) X(96) X(128) X(256) X(512) X(768) X(2304) X(4608)

bool fft2_has_px(int h) { return has_px(h); }
bool fft2_has_pz(int nz, int tile) { return has_pz(nz, tile); }

bool fft2_has_x(int h) {
#define PTC_CASE(N) if (h == N) return true;
  PTC_X_SIZES(PTC_CASE)
#undef PTC_CASE
  return false;
}

int fft2_x_fwd(ptc_ctx *c, const double *src, void *spec, int j0, int nj, bool f64, const DivSrc *div,
               const ResSrc *res) {
  switch (c->pp.h) {
#define PTC_CASE(N) case N: return x_fwd_t<N>(c, src, spec, j0, nj, f64, div, res);
    PTC_X_SIZES(PTC_CASE)
#undef PTC_CASE
    default: return PTC_E_UNSUPPORTED;
  }
}

long long fft2_x_fwd_blocks(const ptc_ctx *c, int nj) {
  switch (c->pp.h) {
#define PTC_CASE(N) case N: return x_fwd_blocks<N>((long long)nj * c->nz);
    PTC_X_SIZES(PTC_CASE)
#undef PTC_CASE
    default: return 0;
  }
}

int fft2_x_inv(ptc_ctx *c, const void *spec, double *dst, double scale, int j0, int nj, bool acc, bool ghost) {
  switch (c->pp.h) {
#define PTC_CASE(N) case N: return x_inv_t<N>(c, spec, dst, scale, j0, nj, acc, ghost);
    PTC_X_SIZES(PTC_CASE)
#undef PTC_CASE
    default: return PTC_E_UNSUPPORTED;
  }
}

// z lengths with a specialised kernel and their column tile (<= 64 KB LDS,
// <= 512 threads): the default tile of each length
#define PTC_Z_SIZES(Z) \
  Z(16, 16) Z(32, 16) Z(64, 16) Z(96, 16) Z(128, 16) Z(192, 16) Z(256, 16) Z(512, 8) Z(700, 4) Z(1400, 2)
// wider alternatives for the autotuner: 256-byte row segments, one 1024-thread
// workgroup of 128 KB LDS per CU; for the long lengths twice the default
// tile, one workgroup of 90 KB LDS per CU
#define PTC_Z_WIDE(Z) Z(256, 32) Z(512, 16) Z(700, 8) Z(1400, 4)

// req = 0: the default tile; else that tile if compiled for nz
bool fft2_has_z(int nz, int req, int *tile) {
#define PTC_CASE(N, T) \
  if (nz == N && (req == 0 || req == T)) { *tile = T; return true; }
  PTC_Z_SIZES(PTC_CASE)
  PTC_Z_WIDE(PTC_CASE)
#undef PTC_CASE
  return false;
}

int fft2_z(ptc_ctx *c, void *spec, bool inverse, int j0, int nj, bool f64) {
  const int T = c->pp.ztile;
#define PTC_CASE(N, TT) \
  if (c->nz == N && T == TT) return z_t<N, TT>(c, spec, inverse, j0, nj, f64);
  PTC_Z_SIZES(PTC_CASE)
  PTC_Z_WIDE(PTC_CASE)
#undef PTC_CASE
  set_error("no specialised z kernel for this length and tile");
  return PTC_E_UNSUPPORTED;
}

}  // namespace ptc
