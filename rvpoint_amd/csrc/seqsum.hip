// seqsum.hip -- the reference's sequential fp32 sum  s = fl(s + x_i), i = 0..n-1, x_i >= 0, computed in
// parallel and bit-exactly (used for SOR's global mean and variance, statistical_outlier_removal.cpp:43-52,
// 193-205).
//
// While the running sum stays in one binade [2^e, 2^(e+1)) it is an integer multiple M of u = 2^(e-23),
// and each step is exact integer arithmetic:  M <- M + round(x_i / u), round-to-nearest-even, where only
// an exact tie (fractional part 1/2) depends on the parity of M.  A step is therefore the map
//   M -> M + (M even ? ie : io),
// and such maps compose associatively:  (ie, io) then (je, jo) = (ie + (ie odd ? jo : je),
//                                                               io + (io odd ? je : jo)).
// The array is cut into chunks of 256.  Per chunk, the binade of the running sum is predicted from fp64
// prefix sums; where the prediction is one binade (with a 1e-3 relative margin), the chunk's map is built
// by an ordered tree reduction over its elements.  One wave then walks the chunks in order: a chunk whose
// map applies (the running sum is in the predicted binade and M stays below 2^24) costs one integer
// step; any other chunk (the first one, binade crossings, zeros, subnormals, non-finite values) is
// replayed element by element with the fp32 adder.  Every applied map is checked, so the result equals
// the sequential sum exactly; the prediction only decides how many chunks take the slow path.
#include "prims.hpp"

namespace rvp {

constexpr int SS_LANE = 4;                  // elements per lane
constexpr int SS_C = WAVE * SS_LANE;        // elements per chunk

struct SSChunk {
  long long ie, io;
  int e;
  int clean;
};

// mode 0: x_i = v[i]; mode 1: x_i = (v[i] - mean)^2 with both ops rounded once (the variance terms)
__device__ __forceinline__ float ss_term(float v, int mode, float mean) {
  if (mode == 0) return v;
  const float d = v - mean;
  return d * d;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int d = WAVE / 2; d > 0; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

__device__ __forceinline__ void ss_compose(long long &fe, long long &fo, long long ge, long long go) {
  const long long ne = fe + ((fe & 1) ? go : ge);
  const long long no = fo + ((fo & 1) ? ge : go);
  fe = ne;
  fo = no;
}

__global__ __launch_bounds__(64) void ss_chunksum_k(const float *__restrict__ v, size_t n, int mode,
                                                     const float *__restrict__ mean_dev, double *__restrict__ csum) {
  const size_t c = blockIdx.x;
  const float mean = mode ? *mean_dev : 0.0f;
  double acc = 0.0;
#pragma unroll
  for (int r = 0; r < SS_LANE; ++r) {
    const size_t i = c * SS_C + (size_t)lane() * SS_LANE + r;
    if (i < n) acc += (double)ss_term(v[i], mode, mean);
  }
  acc = wave_sum_f64(acc);
  if (lane() == 0) csum[c] = acc;
}

__global__ __launch_bounds__(64) void ss_compose_k(const float *__restrict__ v, size_t n, int mode,
                                                    const float *__restrict__ mean_dev,
                                                    const double *__restrict__ csum, SSChunk *__restrict__ out) {
  const size_t c = blockIdx.x;
  const int l = lane();
  double p = 0.0;
  for (size_t j = l; j < c; j += WAVE) p += csum[j];
  p = wave_sum_f64(p);
  const double lo = p, hi = p + csum[c];
  int e = 0;
  bool clean = false;
  if (c > 0 && lo > 0.0 && hi < 1e38) {
    const int elo = ilogb(lo * (1.0 - 1e-3)), ehi = ilogb(hi * (1.0 + 1e-3));
    if (elo == ehi && elo >= -125 && elo <= 126) {
      clean = true;
      e = elo;
    }
  }
  if (!clean) {
    if (l == 0) out[c] = SSChunk{0, 0, 0, 0};
    return;
  }
  const float mean = mode ? *mean_dev : 0.0f;
  const double scale = ldexp(1.0, 23 - e);
  long long fe = 0, fo = 0;
#pragma unroll
  for (int r = 0; r < SS_LANE; ++r) {
    const size_t i = c * SS_C + (size_t)l * SS_LANE + r;
    if (i < n) {
      double t = (double)ss_term(v[i], mode, mean) * scale;  // exact: power-of-two scaling
      if (!(t < 1099511627776.0)) t = 1099511627776.0;       // 2^40: saturate (the chunk then fails its check)
      const double a = floor(t), fr = t - a;
      const long long ai = (long long)a;
      long long ae, ao;
      if (fr > 0.5) {
        ae = ao = ai + 1;
      } else if (fr < 0.5) {
        ae = ao = ai;
      } else {  // exact tie: the result M + a (+1) must be even
        ae = ai + (ai & 1);
        ao = ai + 1 - (ai & 1);
      }
      ss_compose(fe, fo, ae, ao);
    }
  }
  // ordered tree: lane l composes its segment with the one of lane l + d
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const long long ge = __shfl_down(fe, d), go = __shfl_down(fo, d);
    if ((l & (2 * d - 1)) == 0 && l + d < WAVE) ss_compose(fe, fo, ge, go);
  }
  if (l == 0) out[c] = SSChunk{fe, fo, e, 1};
}

// single wave: walk the chunks in order
__global__ __launch_bounds__(64) void ss_drive_k(const float *__restrict__ v, size_t n, int mode,
                                                  const float *__restrict__ mean_dev,
                                                  const SSChunk *__restrict__ ch, size_t nchunks,
                                                  float *__restrict__ out) {
  const int l = lane();
  const float mean = mode ? *mean_dev : 0.0f;
  float s = 0.0f;
  for (size_t c0 = 0; c0 < nchunks; c0 += WAVE) {
    SSChunk mine = (c0 + l < nchunks) ? ch[c0 + l] : SSChunk{0, 0, 0, 0};
    const int cnt = (int)min<size_t>(WAVE, nchunks - c0);
    for (int j = 0; j < cnt; ++j) {
      const size_t c = c0 + j;
      const int clean = __builtin_amdgcn_readlane(mine.clean, j);
      bool ok = false;
      if (clean) {
        const int e = __builtin_amdgcn_readlane(mine.e, j);
        if (s >= ldexpf(1.0f, e) && s < ldexpf(1.0f, e + 1)) {
          const long long ie = (long long)readlane64((uint64_t)mine.ie, j);
          const long long io = (long long)readlane64((uint64_t)mine.io, j);
          const long long M = (long long)ldexpf(s, 23 - e);
          const long long Mn = M + ((M & 1) ? io : ie);
          if (Mn < (1ll << 24)) {
            s = ldexpf((float)Mn, e - 23);
            ok = true;
          }
        }
      }
      if (!ok) {
        float x[SS_LANE];
#pragma unroll
        for (int r = 0; r < SS_LANE; ++r) {
          const size_t i = c * SS_C + (size_t)l * SS_LANE + r;
          x[r] = i < n ? ss_term(v[i], mode, mean) : 0.0f;  // + 0 leaves s unchanged
        }
        for (int q = 0; q < WAVE; ++q) {
#pragma unroll
          for (int r = 0; r < SS_LANE; ++r) s = s + readlanef(x[r], q);
        }
      }
    }
  }
  if (l == 0) *out = s;
}

int exact_seq_sum(const float *v, size_t n, int mode, const float *mean_dev, float *out_dev, hipStream_t s) {
  if (n == 0) {
    RVP_HIP(hipMemsetAsync(out_dev, 0, sizeof(float), s));
    return RVP_OK;
  }
  const size_t nch = (n + SS_C - 1) / SS_C;
  Tmp cs, ch;
  RVP_TRY(cs.alloc(nch * sizeof(double), s));
  RVP_TRY(ch.alloc(nch * sizeof(SSChunk), s));
  RVP_LAUNCH("ss_chunksum_k", s, ss_chunksum_k<<<(unsigned)nch, 64, 0, s>>>(v, n, mode, mean_dev, cs.as<double>()));
  RVP_LAUNCH("ss_compose_k", s, ss_compose_k<<<(unsigned)nch, 64, 0, s>>>(v, n, mode, mean_dev, cs.as<double>(),
                                                                          ch.as<SSChunk>()));
  RVP_LAUNCH("ss_drive_k", s, ss_drive_k<<<1, 64, 0, s>>>(v, n, mode, mean_dev, ch.as<SSChunk>(), nch, out_dev));
  return RVP_OK;
}

}  // namespace rvp

extern "C" int rvp_seq_sum(const float *v, size_t n, int squared_deviation, const float *mean_dev, float *out_dev,
                           rvp_stream_t st) {
  RVP_CHECK(out_dev && (v || n == 0) && (!squared_deviation || mean_dev), RVP_E_ARG, "null");
  return rvp::exact_seq_sum(v, n, squared_deviation ? 1 : 0, mean_dev, out_dev, rvp::S(st));
}
