// Register-resident mixed-radix complex FFT for small lengths (gfx950).
//
// One thread transforms one line of N complex values held in a register
// array. All loops are unrolled at compile time (sfor), every twiddle is a
// compile-time constant computed in double precision with exact octant
// reduction (so W^0, W^{N/4}, W^{N/2}, ... are exact), and multiplications
// by 1, -1, +-i and the 1/8-turn twiddles are specialised away.
//
// Decimation in time: N = R * M, the R interleaved sub-sequences x[m R + r]
// are transformed (length M), multiplied by W_N^{r k} and recombined by an
// R-point DFT. Radix preference: 4, then 2, 3, 5, 7, then a direct odd DFT.
#pragma once

#include <hip/hip_runtime.h>

namespace fia {

template <int I>
struct IC {
  static constexpr int value = I;
};

template <int B, int E, class F>
__host__ __device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(IC<B>{});
    sfor<B + 1, E>(f);
  }
}

// ---------------------------------------------------------------- constexpr trig
struct dcplx {
  double re, im;
};

constexpr double kHalfPi = 1.57079632679489661923132169163975144;

constexpr double poly_sin(double x) {  // |x| <= pi/4
  double x2 = x * x, term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term *= -x2 / double((2 * k) * (2 * k + 1));
    sum += term;
  }
  return sum;
}

constexpr double poly_cos(double x) {  // |x| <= pi/4
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) {
    term *= -x2 / double((2 * k - 1) * (2 * k));
    sum += term;
  }
  return sum;
}

// exp(2 pi i j / n) for integers j, n > 0
constexpr dcplx cexp_frac(long j, long n) {
  j %= n;
  if (j < 0) j += n;
  long q = (8 * j + n) / (2 * n);  // nearest quarter turn
  long num = 4 * j - q * n;         // |num| <= n / 2
  double r = kHalfPi * double(num) / double(n);
  double c = poly_cos(r), s = poly_sin(r);
  switch (q & 3) {
    case 0: return {c, s};
    case 1: return {-s, c};
    case 2: return {-c, -s};
    default: return {s, -c};
  }
}

constexpr int pick_radix(int n) {
  if (n % 4 == 0 && n != 4) return 4;
  if (n % 2 == 0 && n != 2) return 2;
  if (n % 3 == 0 && n != 3) return 3;
  if (n % 5 == 0 && n != 5) return 5;
  if (n % 7 == 0 && n != 7) return 7;
  for (int p = 11; p * p <= n; p += 2)
    if (n % p == 0) return p;
  return n;  // prime (or 1..4): direct transform
}

// ---------------------------------------------------------------- complex helpers
__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by s * i (s = +1 or -1)
template <int S>
__host__ __device__ __forceinline__ float2 mul_si(float2 a) {
  if constexpr (S > 0) return make_float2(-a.y, a.x);
  else return make_float2(a.y, -a.x);
}

// v * exp(DIR 2 pi i J / N), specialised on the constant
template <int J, int N, int DIR>
__host__ __device__ __forceinline__ float2 twmul(float2 v) {
  constexpr dcplx w = cexp_frac(long(DIR) * J, N);
  constexpr float wr = float(w.re), wi = float(w.im);
  if constexpr (wr == 1.0f && wi == 0.0f) return v;
  else if constexpr (wr == -1.0f && wi == 0.0f) return make_float2(-v.x, -v.y);
  else if constexpr (wr == 0.0f && wi == 1.0f) return make_float2(-v.y, v.x);
  else if constexpr (wr == 0.0f && wi == -1.0f) return make_float2(v.y, -v.x);
  else if constexpr (wr == wi) return make_float2(wr * (v.x - v.y), wr * (v.x + v.y));
  else if constexpr (wr == -wi) return make_float2(wr * (v.x + v.y), wr * (v.y - v.x));
  else return make_float2(v.x * wr - v.y * wi, v.x * wi + v.y * wr);
}

// ---------------------------------------------------------------- base DFTs
template <int DIR>
__host__ __device__ __forceinline__ void dft2(float2& a, float2& b) {
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int DIR>
__host__ __device__ __forceinline__ void dft4(float2 (&x)[4]) {
  float2 t0 = cadd(x[0], x[2]);
  float2 t1 = csub(x[0], x[2]);
  float2 t2 = cadd(x[1], x[3]);
  float2 t3 = mul_si<DIR>(csub(x[1], x[3]));
  x[0] = cadd(t0, t2);
  x[2] = csub(t0, t2);
  x[1] = cadd(t1, t3);
  x[3] = csub(t1, t3);
}

// direct DFT of odd length P, pairing r and P - r
template <int P, int DIR>
__host__ __device__ __forceinline__ void dft_odd(float2 (&x)[P]) {
  constexpr int H = (P - 1) / 2;
  float2 s[H + 1], d[H + 1];
  float2 x0 = x[0];
  float2 sum = x0;
  sfor<1, H + 1>([&](auto r_) {
    constexpr int r = decltype(r_)::value;
    s[r] = cadd(x[r], x[P - r]);
    d[r] = csub(x[r], x[P - r]);
    sum = cadd(sum, s[r]);
  });
  float2 out[P];
  out[0] = sum;
  sfor<1, P>([&](auto q_) {
    constexpr int q = decltype(q_)::value;
    float2 acc_c = x0;
    float2 acc_s = make_float2(0.f, 0.f);
    sfor<1, H + 1>([&](auto r_) {
      constexpr int r = decltype(r_)::value;
      constexpr dcplx w = cexp_frac(long(r) * q, P);
      constexpr float c = float(w.re), sn = float(w.im);
      acc_c.x += s[r].x * c;
      acc_c.y += s[r].y * c;
      acc_s.x += d[r].x * sn;
      acc_s.y += d[r].y * sn;
    });
    out[q] = cadd(acc_c, mul_si<DIR>(acc_s));
  });
  sfor<0, P>([&](auto q_) {
    constexpr int q = decltype(q_)::value;
    x[q] = out[q];
  });
}

// ---------------------------------------------------------------- recursive FFT
// In place: x[k] <- sum_j x[j] exp(DIR 2 pi i j k / N), unnormalised.
template <int N, int DIR>
__host__ __device__ __forceinline__ void fft(float2 (&x)[N]) {
  if constexpr (N == 1) {
    return;
  } else if constexpr (N == 2) {
    dft2<DIR>(x[0], x[1]);
  } else if constexpr (N == 4) {
    dft4<DIR>(x);
  } else if constexpr (pick_radix(N) == N) {
    dft_odd<N, DIR>(x);
  } else {
    constexpr int R = pick_radix(N);
    constexpr int M = N / R;
    float2 sub[R][M];
    sfor<0, R>([&](auto r_) {
      constexpr int r = decltype(r_)::value;
      sfor<0, M>([&](auto m_) {
        constexpr int m = decltype(m_)::value;
        sub[r][m] = x[m * R + r];
      });
      fft<M, DIR>(sub[r]);
    });
    sfor<0, M>([&](auto k_) {
      constexpr int k = decltype(k_)::value;
      float2 t[R];
      sfor<0, R>([&](auto r_) {
        constexpr int r = decltype(r_)::value;
        t[r] = twmul<r * k, N, DIR>(sub[r][k]);
      });
      fft<R, DIR>(t);
      sfor<0, R>([&](auto q_) {
        constexpr int q = decltype(q_)::value;
        x[k + q * M] = t[q];
      });
    });
  }
}

// ---------------------------------------------------------------- in-place FFTs
// The same transform on x[O + S j], j = 0..N-1, without copies (one array of
// N values plus R temporaries live). fft_dif takes the input in natural
// order and leaves X[k] at j = fft_pos(N, k); fft_dit takes x[n] at
// j = fft_pos(N, n) and leaves X[k] at j = k. With N = R M (R =
// pick_radix(N)): fft_pos(N, q + R k') = q M + fft_pos(M, k'), the identity
// for the base lengths (2, 4, primes).
__host__ __device__ constexpr int fft_pos(int n, int k) {
  if (n <= 4 || pick_radix(n) == n) return k;
  return (k % pick_radix(n)) * (n / pick_radix(n)) + fft_pos(n / pick_radix(n), k / pick_radix(n));
}

template <int N, int DIR, int O, int S, int L>
__host__ __device__ __forceinline__ void fft_base_at(float2 (&x)[L]) {
  float2 t[N];
  sfor<0, N>([&](auto j_) {
    constexpr int j = decltype(j_)::value;
    t[j] = x[O + S * j];
  });
  fft<N, DIR>(t);
  sfor<0, N>([&](auto j_) {
    constexpr int j = decltype(j_)::value;
    x[O + S * j] = t[j];
  });
}

// decimation in frequency: R-point DFTs over x[m + r M], twiddles
// W_N^{q m}, then the M-point transforms of the R blocks q M + m
template <int N, int DIR, int O, int S, int L>
__host__ __device__ __forceinline__ void fft_dif(float2 (&x)[L]) {
  if constexpr (N == 1) {
    return;
  } else if constexpr (N <= 4 || pick_radix(N) == N) {
    fft_base_at<N, DIR, O, S, L>(x);
  } else {
    constexpr int R = pick_radix(N);
    constexpr int M = N / R;
    sfor<0, M>([&](auto m_) {
      constexpr int m = decltype(m_)::value;
      float2 t[R];
      sfor<0, R>([&](auto r_) {
        constexpr int r = decltype(r_)::value;
        t[r] = x[O + S * (m + r * M)];
      });
      fft<R, DIR>(t);
      sfor<0, R>([&](auto q_) {
        constexpr int q = decltype(q_)::value;
        x[O + S * (m + q * M)] = twmul<q * m, N, DIR>(t[q]);
      });
    });
    sfor<0, R>([&](auto q_) {
      constexpr int q = decltype(q_)::value;
      fft_dif<M, DIR, O + S * q * M, S, L>(x);
    });
  }
}

// decimation in time: the M-point transforms of the R blocks (block q holds
// the sub-sequence x[m R + q]), then twiddles W_N^{q k} and R-point DFTs
// over the blocks, X[k + r M] at k + r M
template <int N, int DIR, int O, int S, int L>
__host__ __device__ __forceinline__ void fft_dit(float2 (&x)[L]) {
  if constexpr (N == 1) {
    return;
  } else if constexpr (N <= 4 || pick_radix(N) == N) {
    fft_base_at<N, DIR, O, S, L>(x);
  } else {
    constexpr int R = pick_radix(N);
    constexpr int M = N / R;
    sfor<0, R>([&](auto q_) {
      constexpr int q = decltype(q_)::value;
      fft_dit<M, DIR, O + S * q * M, S, L>(x);
    });
    sfor<0, M>([&](auto k_) {
      constexpr int k = decltype(k_)::value;
      float2 t[R];
      sfor<0, R>([&](auto q_) {
        constexpr int q = decltype(q_)::value;
        t[q] = twmul<q * k, N, DIR>(x[O + S * (q * M + k)]);
      });
      fft<R, DIR>(t);
      sfor<0, R>([&](auto r_) {
        constexpr int r = decltype(r_)::value;
        x[O + S * (r * M + k)] = t[r];
      });
    });
  }
}

}  // namespace fia
