// Shared device helpers for the MI355X (gfx950) drift-registration kernels:
// complex arithmetic on float2/double2, compile-time twiddle constants,
// in-register small DFTs, an LDS Stockham FFT for one workgroup, and the
// XCD-aware block-index remap.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#define ZR_DEV __device__ __forceinline__

// ---------------------------------------------------------------- complex
ZR_DEV float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
ZR_DEV float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
ZR_DEV float2 c_mul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// conj(a) * b
ZR_DEV float2 c_cjmul(float2 a, float2 b) { return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }
ZR_DEV float2 c_scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
ZR_DEV float2 c_conj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by DIR*i  (DIR = -1 forward, +1 inverse)
template <int DIR> ZR_DEV float2 c_muli(float2 a) {
  return DIR > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ------------------------------------------------- compile-time sin/cos
// Own Taylor-series evaluation (argument reduced to [-pi/4, pi/4]) so twiddle
// constants of the in-register DFTs are exact to double precision at compile time.
namespace zr_ct {
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double sin_series(double x) {
  double t = x, s = x;
  for (int k = 1; k < 14; ++k) { t *= -x * x / ((2 * k) * (2 * k + 1)); s += t; }
  return s;
}
constexpr double cos_series(double x) {
  double t = 1.0, s = 1.0;
  for (int k = 1; k < 14; ++k) { t *= -x * x / ((2 * k - 1) * (2 * k)); s += t; }
  return s;
}
// cos/sin of 2*pi*num/den for integers
constexpr double cos2pi(long num, long den) {
  long n = ((num % den) + den) % den;           // [0, den)
  long q8 = (8 * n) / den;                       // octant
  double x = 2.0 * kPi * (double)n / (double)den;
  // reduce to [-pi/4, pi/4] around the nearest multiple of pi/2
  long quad = (q8 + 1) / 2;                      // 0..4
  double r = x - quad * (kPi / 2.0);
  switch (quad & 3) {
    case 0: return cos_series(r);
    case 1: return -sin_series(r);
    case 2: return -cos_series(r);
    default: return sin_series(r);
  }
}
constexpr double sin2pi(long num, long den) { return cos2pi(num * 4 - den, 4 * den); }
}  // namespace zr_ct

template <int DIR, int N, int K> ZR_DEV float2 ct_twiddle() {
  // exp(DIR * 2*pi*i*K/N)
  constexpr float c = (float)zr_ct::cos2pi(K, N);
  constexpr float s = (float)(DIR * zr_ct::sin2pi(K, N));
  return make_float2(c, s);
}

// -------------------------------------------------- in-register DFTs
// X[k] = sum_n x[n] exp(DIR*2*pi*i*n*k/R), in place, natural order in/out.
template <int R, int DIR> struct Dft;

template <int DIR> struct Dft<1, DIR> { static ZR_DEV void run(float2*) {} };

template <int DIR> struct Dft<2, DIR> {
  static ZR_DEV void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = c_add(a, b);
    v[1] = c_sub(a, b);
  }
};

template <int DIR> struct Dft<3, DIR> {
  static ZR_DEV void run(float2* v) {
    constexpr float h = -0.5f;
    constexpr float s = 0.86602540378443864676f;  // sqrt(3)/2
    float2 a = v[0], b = v[1], c = v[2];
    float2 t1 = c_add(b, c);
    float2 t2 = make_float2(a.x + h * t1.x, a.y + h * t1.y);
    float2 d = c_sub(b, c);
    float2 t3 = c_muli<DIR>(make_float2(s * d.x, s * d.y));
    v[0] = c_add(a, t1);
    v[1] = c_add(t2, t3);
    v[2] = c_sub(t2, t3);
  }
};

template <int DIR> struct Dft<4, DIR> {
  static ZR_DEV void run(float2* v) {
    float2 a = v[0], b = v[1], c = v[2], d = v[3];
    float2 s0 = c_add(a, c), s1 = c_sub(a, c);
    float2 s2 = c_add(b, d), s3 = c_muli<DIR>(c_sub(b, d));
    v[0] = c_add(s0, s2);
    v[2] = c_sub(s0, s2);
    v[1] = c_add(s1, s3);
    v[3] = c_sub(s1, s3);
  }
};

template <int DIR> struct Dft<5, DIR> {
  static ZR_DEV void run(float2* v) {
    constexpr float c1 = (float)zr_ct::cos2pi(1, 5), c2 = (float)zr_ct::cos2pi(2, 5);
    constexpr float s1 = (float)zr_ct::sin2pi(1, 5), s2 = (float)zr_ct::sin2pi(2, 5);
    float2 a = v[0];
    float2 p1 = c_add(v[1], v[4]), m1 = c_sub(v[1], v[4]);
    float2 p2 = c_add(v[2], v[3]), m2 = c_sub(v[2], v[3]);
    float2 r1 = make_float2(a.x + c1 * p1.x + c2 * p2.x, a.y + c1 * p1.y + c2 * p2.y);
    float2 r2 = make_float2(a.x + c2 * p1.x + c1 * p2.x, a.y + c2 * p1.y + c1 * p2.y);
    float2 i1 = c_muli<DIR>(make_float2(s1 * m1.x + s2 * m2.x, s1 * m1.y + s2 * m2.y));
    float2 i2 = c_muli<DIR>(make_float2(s2 * m1.x - s1 * m2.x, s2 * m1.y - s1 * m2.y));
    v[0] = c_add(a, c_add(p1, p2));
    v[1] = c_add(r1, i1);
    v[4] = c_sub(r1, i1);
    v[2] = c_add(r2, i2);
    v[3] = c_sub(r2, i2);
  }
};

// Composite R = R1*R2 (Cooley-Tukey in registers):
// n = R2*n1 + n2, k = k1 + R1*k2.
template <int R1, int R2, int DIR> struct DftCT {
  static ZR_DEV void run(float2* v) {
    constexpr int R = R1 * R2;
    float2 t[R];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      float2 u[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) u[n1] = v[R2 * n1 + n2];
      Dft<R1, DIR>::run(u);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) t[n2 * R1 + k1] = u[k1];
    }
#pragma unroll
    for (int n2 = 1; n2 < R2; ++n2)
#pragma unroll
      for (int k1 = 1; k1 < R1; ++k1) t[n2 * R1 + k1] = mul_tw(t[n2 * R1 + k1], n2 * k1);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      float2 u[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) u[n2] = t[n2 * R1 + k1];
      Dft<R2, DIR>::run(u);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = u[k2];
    }
  }
  // a * exp(DIR*2*pi*i*j/R): a quarter turn (4 j a multiple of R) as a swap and
  // sign change (a product with the constants 0 and +-1 is not folded without
  // fast math), otherwise a complex product with the constant
  static ZR_DEV float2 mul_tw(float2 a, int j) {
    constexpr int R = R1 * R2;
    if ((4 * j) % R == 0) {
      const int q = ((4 * j) / R) % 4;
      if (q == 0) return a;
      if (q == 2) return make_float2(-a.x, -a.y);
      const float2 b = c_muli<DIR>(a);          // a * (DIR i)
      return q == 1 ? b : make_float2(-b.x, -b.y);
    }
    return c_mul(a, tw_ct(j));
  }
  static ZR_DEV float2 tw_ct(int j) {
    // exp(DIR*2*pi*i*j/R) from a compile-time table (j is a constant after unrolling)
    constexpr int R = R1 * R2;
    float c = 1.f, s = 0.f;
    switch (j % R) {
#define ZR_TW_CASE(K) case K: { c = (float)zr_ct::cos2pi(K, R); s = (float)(DIR * zr_ct::sin2pi(K, R)); } break;
      ZR_TW_CASE(0) ZR_TW_CASE(1) ZR_TW_CASE(2) ZR_TW_CASE(3) ZR_TW_CASE(4) ZR_TW_CASE(5) ZR_TW_CASE(6)
      ZR_TW_CASE(7) ZR_TW_CASE(8) ZR_TW_CASE(9) ZR_TW_CASE(10) ZR_TW_CASE(11) ZR_TW_CASE(12) ZR_TW_CASE(13)
      ZR_TW_CASE(14) ZR_TW_CASE(15) ZR_TW_CASE(16) ZR_TW_CASE(17) ZR_TW_CASE(18) ZR_TW_CASE(19)
      ZR_TW_CASE(20) ZR_TW_CASE(21) ZR_TW_CASE(22) ZR_TW_CASE(23) ZR_TW_CASE(24) ZR_TW_CASE(25)
      ZR_TW_CASE(26) ZR_TW_CASE(27) ZR_TW_CASE(28) ZR_TW_CASE(29) ZR_TW_CASE(30) ZR_TW_CASE(31)
      ZR_TW_CASE(32) ZR_TW_CASE(33) ZR_TW_CASE(34) ZR_TW_CASE(35) ZR_TW_CASE(36) ZR_TW_CASE(37)
      ZR_TW_CASE(38) ZR_TW_CASE(39) ZR_TW_CASE(40) ZR_TW_CASE(41) ZR_TW_CASE(42) ZR_TW_CASE(43)
      ZR_TW_CASE(44) ZR_TW_CASE(45) ZR_TW_CASE(46) ZR_TW_CASE(47) ZR_TW_CASE(48) ZR_TW_CASE(49)
      ZR_TW_CASE(50) ZR_TW_CASE(51) ZR_TW_CASE(52) ZR_TW_CASE(53) ZR_TW_CASE(54) ZR_TW_CASE(55)
      ZR_TW_CASE(56) ZR_TW_CASE(57) ZR_TW_CASE(58) ZR_TW_CASE(59) ZR_TW_CASE(60) ZR_TW_CASE(61)
      ZR_TW_CASE(62) ZR_TW_CASE(63)
#undef ZR_TW_CASE
      default: break;
    }
    return make_float2(c, s);
  }
};

template <int DIR> struct Dft<8, DIR> { static ZR_DEV void run(float2* v) { DftCT<2, 4, DIR>::run(v); } };
template <int DIR> struct Dft<16, DIR> { static ZR_DEV void run(float2* v) { DftCT<4, 4, DIR>::run(v); } };
template <int DIR> struct Dft<32, DIR> { static ZR_DEV void run(float2* v) { DftCT<2, 16, DIR>::run(v); } };
template <int DIR> struct Dft<48, DIR> { static ZR_DEV void run(float2* v) { DftCT<3, 16, DIR>::run(v); } };
template <int DIR> struct Dft<64, DIR> { static ZR_DEV void run(float2* v) { DftCT<4, 16, DIR>::run(v); } };
template <int DIR> struct Dft<24, DIR> { static ZR_DEV void run(float2* v) { DftCT<3, 8, DIR>::run(v); } };

// ------------------------------------------------- radix plan (compile time)
constexpr int zr_pow2_part(int n) { int p = 1; while (n % 2 == 0) { n /= 2; p *= 2; } return p; }
constexpr int zr_next_radix(int rem) {
  int p2 = zr_pow2_part(rem);
  if (p2 == 32) return 8;
  if (p2 >= 16) return 16;
  if (p2 == 8) return 8;
  if (p2 == 4) return 4;
  if (p2 == 2) return 2;
  if (rem % 3 == 0) return 3;
  if (rem % 5 == 0) return 5;
  return 1;  // unsupported factor (compile-time guarded by static_assert below)
}
constexpr bool zr_smooth(int n) {
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  return n == 1;
}

// LDS layout of the passes between the first and the last: one padding
// element after every 16, so that the radix-R outputs a thread writes side by
// side (R consecutive float2) land in different banks for neighbouring lanes.
// The first pass reads the caller's natural order and the last pass writes it
// back, so callers fill and read buf[0..L) naturally; the buffer must hold
// zr_fft_lds(L) elements.
ZR_DEV int zr_pad16(int i) { return i + (i >> 4); }
constexpr int zr_fft_lds(int L) { return L + L / 16 + 1; }

// Workgroup barrier that orders LDS only: the fences are scoped to the local
// address space, so global loads still in flight (a prefetch of the next
// input) and stores are not waited for, as __syncthreads() would.
ZR_DEV void zr_lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Twiddle sources of the LDS FFT: get(t) = exp(-2*pi*i*t/L) for t in [0, L).
// GlobalTw reads a full table in global memory (one load per twiddle, from L2);
// LdsTw forms it from two short tables staged in LDS, hi[t / 64] * lo[t % 64]
// (two LDS reads and a complex product: no global-memory latency inside the
// passes; within about one float32 ulp of the full table).
struct GlobalTw {
  const float2* __restrict__ tw;
  ZR_DEV float2 get(int t) const { return tw[t]; }
};
struct LdsTw {
  const float2* hi;   // exp(-2 pi i 64 h / L), h < L / 64
  const float2* lo;   // exp(-2 pi i l / L), l < 64
  ZR_DEV float2 get(int t) const { return c_mul(hi[t >> 6], lo[t & 63]); }
};
// LDS floats2 an LdsTw table for length L takes.
constexpr int zr_ldstw_size(int L) { return (L + 63) / 64 + 64; }
// Fill an LdsTw table (hi first, then lo) from the full global table tw of
// length L; a block's threads 0 .. nthreads-1 share the work.  The caller
// synchronises the workgroup before the first use.
ZR_DEV LdsTw zr_ldstw_fill(float2* lds, const float2* __restrict__ tw, int L, int tid, int nthreads) {
  const int nh = (L + 63) / 64;
  for (int e = tid; e < nh + 64; e += nthreads) lds[e] = e < nh ? tw[64 * e] : (e - nh < L ? tw[e - nh] : make_float2(1.f, 0.f));
  return LdsTw{lds, lds + nh};
}

// One Stockham pass of radix R with stride Ls over an LDS-resident sequence of
// length L, in place (all reads of the pass complete before any write).
// tws.get(t) = exp(-2*pi*i*t/L), t in [0,L) (forward); DIR=+1 uses conj.
// INP / OUTP: the pass reads / writes the padded layout.
template <int L, int R, int Ls, int DIR, int NT, bool INP, bool OUTP, class TW>
ZR_DEV void stockham_pass(float2* buf, const TW& tws, int tid) {
  constexpr int NB = L / R;                       // butterflies
  constexpr int PER = (NB + NT - 1) / NT;         // per thread
  float2 v[PER][R];
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    int j = tid + b * NT;
    if (NB % NT == 0 || j < NB) {
      int k = j % Ls;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = j + r * NB;
        v[b][r] = buf[INP ? zr_pad16(i) : i];
      }
      if (Ls > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tws.get((k * r) * (L / (Ls * R)));
          if (DIR > 0) w.y = -w.y;
          v[b][r] = c_mul(v[b][r], w);
        }
      }
      Dft<R, DIR>::run(v[b]);
    }
  }
  zr_lds_barrier();
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    int j = tid + b * NT;
    if (NB % NT == 0 || j < NB) {
      int k = j % Ls;
      int base = (j / Ls) * Ls * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = base + r * Ls;
        buf[OUTP ? zr_pad16(i) : i] = v[b][r];
      }
    }
  }
  zr_lds_barrier();
}

template <int L, int Ls, int DIR, int NT> struct WgFft {
  template <class TW>
  static ZR_DEV void run(float2* buf, const TW& tws, int tid) {
    constexpr int R = zr_next_radix(L / Ls);
    static_assert(R > 1, "FFT length must be 2^a 3^b 5^c");
    constexpr bool first = (Ls == 1), last = (Ls * R == L);
    stockham_pass<L, R, Ls, DIR, NT, !first, !last>(buf, tws, tid);
    WgFft<L, Ls * R, DIR, NT>::run(buf, tws, tid);
  }
};
template <int L, int DIR, int NT> struct WgFft<L, L, DIR, NT> {
  template <class TW>
  static ZR_DEV void run(float2*, const TW&, int) {}
};

// First pass of wg_fft taken from registers: when the first radix R =
// zr_next_radix(L) has one butterfly per thread (L / R == NT), thread tid
// holds x[r] = element tid + r (L / R), r < R (the pass's own inputs), so the
// fill of buf and the pass's reads are skipped.  The pass output goes to buf
// in the layout of the following passes; the caller synchronises the
// workgroup (zr_lds_barrier), then runs wg_fft_rest.
template <int L, int NT> constexpr bool wg_fft_regs_first() { return L / zr_next_radix(L) == NT; }
template <int L, int DIR, int NT>
ZR_DEV void wg_fft_first_regs(float2 (&x)[zr_next_radix(L)], float2* buf, int tid) {
  constexpr int R = zr_next_radix(L);
  static_assert(L / R == NT, "one first-pass butterfly per thread");
  Dft<R, DIR>::run(x);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = tid * R + r;
    buf[R == L ? i : zr_pad16(i)] = x[r];
  }
}
template <int L, int DIR, int NT>
ZR_DEV void wg_fft_rest(float2* buf, const LdsTw& tws, int tid) {
  WgFft<L, zr_next_radix(L), DIR, NT>::run(buf, tws, tid);
}

// Whole-workgroup complex FFT of length L held in LDS (natural order in/out).
// buf holds zr_fft_lds(L) elements (the passes in between use the padding).
// The caller must synchronise the workgroup after filling buf; returns
// synchronised (for LDS: its barriers do not wait for global memory).
// Twiddles from the full global table tw of length L.
template <int L, int DIR, int NT>
ZR_DEV void wg_fft(float2* buf, const float2* __restrict__ tw, int tid) {
  static_assert(zr_smooth(L), "FFT length must be 2^a 3^b 5^c");
  WgFft<L, 1, DIR, NT>::run(buf, GlobalTw{tw}, tid);
}
// The same transform with twiddles from an LDS table (zr_ldstw_fill).
template <int L, int DIR, int NT>
ZR_DEV void wg_fft(float2* buf, const LdsTw& tws, int tid) {
  static_assert(zr_smooth(L), "FFT length must be 2^a 3^b 5^c");
  WgFft<L, 1, DIR, NT>::run(buf, tws, tid);
}

// ---------------------------------------------------------- XCD remap
// Workgroups are dispatched round-robin over the 8 XCDs (id % 8).  Remap so
// that consecutive logical blocks run on the same XCD and share its L2.
ZR_DEV int xcd_remap(int id, int nblocks) {
  constexpr int X = 8;
  int q = nblocks / X, r = nblocks % X;
  int xcd = id % X, k = id / X;
  return (xcd < r) ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

// ----------------------------------------------------------- reductions
ZR_DEV double wave_sum_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// x + x[lane ^ 16] and x + x[lane ^ 32] in every lane, through the gfx950
// row-swap permutes (VALU ops: no LDS round trip as with a bpermute).  The
// swap leaves a lane's own value in one result and its partner's in the other,
// so the sum is own + partner (an IEEE add is commutative: the same bits as
// x + shfl_xor(x)).
ZR_DEV float xor16_sum(float x) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
ZR_DEV float xor32_sum(float x) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
ZR_DEV float wave_sum_f(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ------------------------------------------------------ host: LDS opt-in
// Let a kernel take all of a workgroup's LDS (160 KiB on gfx950, minus its
// static LDS) as dynamic memory, once per process.  The dynamic size a launch
// asks for varies with the movie's shape (maxShift, Gaussian radius, frame
// count), and movies may launch from several threads; one maximal opt-in
// avoids racing per-size updates.
inline void zr_allow_full_lds(const void* fn, std::once_flag& flag) {
  std::call_once(flag, [fn] {
    hipFuncAttributes fa{};
    size_t stat = 0;
    if (hipFuncGetAttributes(&fa, fn) == hipSuccess) stat = fa.sharedSizeBytes;
    const size_t lim = 160 * 1024;
    hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(lim > stat ? lim - stat : 0));
  });
}
