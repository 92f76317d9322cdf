// seqsum.hip -- the reference's sequential fp32 sum  s = fl(s + x_i), i = 0..n-1, x_i >= 0, computed in
// parallel and bit-exactly (used for SOR's global mean and variance, statistical_outlier_removal.cpp:43-52,
// 193-205), for one array or for many independent segments (one per cloud of a batch) at once.
//
// While the running sum stays in one binade [2^e, 2^(e+1)) it is an integer multiple M of u = 2^(e-23),
// and each step is exact integer arithmetic:  M <- M + round(x_i / u), round-to-nearest-even, where only
// an exact tie (fractional part 1/2) depends on the parity of M.  A step is therefore the map
//   M -> M + (M even ? ie : io),
// and such maps compose associatively:  (ie, io) then (je, jo) = (ie + (ie odd ? jo : je),
//                                                               io + (io odd ? je : jo)).
// Each segment is cut into chunks of 256.  Per chunk, the binade of the running sum is predicted from fp64
// prefix sums; where the prediction is one binade (with a 1e-3 relative margin), the chunk's map is built
// by an ordered tree reduction over its elements.  One wave per segment then walks its chunks in order: a
// chunk whose map applies (the running sum is in the predicted binade and M stays below 2^24) costs one
// integer step; any other chunk (the first one, binade crossings, zeros, subnormals, non-finite values)
// is replayed: lane maps (SS_LANE terms each) in the running sum's binade, up to the lane where it leaves
// the binade, which the fp32 adder replays term by term.  Every applied map is checked, so the result
// equals the sequential sum exactly; the prediction only decides how many chunks take the slow path.
#include "prims.hpp"

namespace rvp {

constexpr int SS_LANE = 4;                  // elements per lane
constexpr int SS_C = WAVE * SS_LANE;        // elements per chunk

// Maps are kept in 32 bits, saturated at SS_SAT = 2^25: a composed value whose exact counterpart is below
// 2^24 is exact (its inputs are, so the parity selections agree), and one whose exact counterpart is 2^24
// or more stays at 2^24 or more.  A state M0 + map (M0 < 2^24) then fails the check M < 2^24 exactly when
// the exact one does.
constexpr int SS_SAT = 1 << 25;

struct SSChunk {
  int ie, io;
  int e;
  int clean;
};

// mode 0: x_i = v[i]; modes 1, 2: x_i = (v[i] - mean)^2 with both ops rounded once (the variance terms)
__device__ __forceinline__ float ss_term(float v, int mode, float mean) {
  if (mode == 0) return v;
  const float d = v - mean;
  return d * d;
}

// the segment's mean: mode 1 reads it, mode 2 derives it from the segment's sum as fl(sum / len)
__device__ __forceinline__ float ss_mean(int mode, const float *__restrict__ mean_dev,
                                         const uint32_t *__restrict__ seg, int s) {
  if (mode == 0) return 0.0f;
  if (mode == 1) return mean_dev[s];
  return __fdiv_rn(mean_dev[s], __uint2float_rn(seg[s + 1] - seg[s]));
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int d = WAVE / 2; d > 0; d >>= 1) v += shfl_xor(v, d);
  return v;
}

__device__ __forceinline__ void ss_compose(int &fe, int &fo, int ge, int go) {
  const int ne = min(fe + ((fe & 1) ? go : ge), SS_SAT);
  const int no = min(fo + ((fo & 1) ? ge : go), SS_SAT);
  fe = ne;
  fo = no;
}

// The step map (ae, ao) of one term x in binade e, k = 23 - e: M -> M + (M even ? ae : ao).  A term that
// is negative, non-finite or of SS_SAT units or more gets the saturated map, so the state check (M < 2^24)
// fails and the term is added by the fp32 adder instead.  t = x * 2^k is exact unless it underflows, and
// an underflowed t is below 2^-126, which maps to 0 like the exact one; floor and t - floor(t) are exact.
__device__ __forceinline__ void ss_term_map(float x, int k, int &ae, int &ao) {
  float t = ldexpf(x, k);
  if (!(t >= 0.0f && t < (float)SS_SAT)) t = (float)SS_SAT;
  const float a = floorf(t), fr = t - a;
  const int ai = (int)a;
  if (fr > 0.5f) {
    ae = ao = ai + 1;
  } else if (fr < 0.5f) {
    ae = ao = ai;
  } else {  // exact tie: the result M + a (+1) must be even
    ae = ai + (ai & 1);
    ao = ai + 1 - (ai & 1);
  }
}

// one scan step: compose the map of the lane the DPP control selects (identity (0, 0) where it selects
// none or the row is masked off) in front of this lane's
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void ss_scan_step(int &fe, int &fo) {
  int pe = __builtin_amdgcn_update_dpp(0, fe, CTRL, ROW_MASK, 0xF, true);
  int po = __builtin_amdgcn_update_dpp(0, fo, CTRL, ROW_MASK, 0xF, true);
  ss_compose(pe, po, fe, fo);
  fe = pe;
  fo = po;
}

// inclusive scan of step maps over the wave, earlier lanes' maps first ((0, 0) is the identity): within
// rows of 16 lanes by row_shr 1, 2, 4, 8, then lane 15 of rows 0 and 2 into rows 1 and 3 (row_bcast:15),
// then lane 31 into rows 2 and 3 (row_bcast:31).  The walk over the chunks is one wave, so the scan's
// latency is its cost: DPP moves instead of LDS permutes.
__device__ __forceinline__ void ss_scan_maps(int &fe, int &fo) {
  ss_scan_step<0x111, 0xF>(fe, fo);
  ss_scan_step<0x112, 0xF>(fe, fo);
  ss_scan_step<0x114, 0xF>(fe, fo);
  ss_scan_step<0x118, 0xF>(fe, fo);
  ss_scan_step<0x142, 0xA>(fe, fo);
  ss_scan_step<0x143, 0xC>(fe, fo);
}

// largest s in [0, nseg) with off[s] <= v (off non-decreasing, off[0] == 0)
__device__ __forceinline__ int seg_of(const uint32_t *__restrict__ off, int nseg, uint32_t v) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= v) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void ss_nchunks_k(const uint32_t *__restrict__ seg, int nseg, uint32_t *__restrict__ nch) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nseg) nch[s] = (seg[s + 1] - seg[s] + SS_C - 1) / SS_C;
}

// the same counts and their exclusive scan (coff[nseg] = total) in one wave, for few segments
__global__ __launch_bounds__(64) void ss_chunk_offsets_k(const uint32_t *__restrict__ seg, int nseg,
                                                          uint32_t *__restrict__ coff) {
  const int l = lane();
  uint32_t carry = 0;
  for (int b0 = 0; b0 < nseg; b0 += WAVE) {
    const int s = b0 + l;
    const uint32_t c = s < nseg ? (seg[s + 1] - seg[s] + SS_C - 1) / SS_C : 0u;
    const uint32_t inc = wave_incl_scan(c);
    if (s < nseg) coff[s] = carry + inc - c;
    carry += (uint32_t)__builtin_amdgcn_readlane((int)inc, WAVE - 1);
  }
  if (l == 0) coff[nseg] = carry;
}

struct ChunkRef {
  int s;          // segment
  uint32_t a, b;  // element range
  uint32_t c0;    // first chunk of the segment
};

__device__ __forceinline__ bool chunk_ref(const uint32_t *__restrict__ seg, const uint32_t *__restrict__ coff,
                                          int nseg, uint32_t c, ChunkRef &r) {
  if (c >= coff[nseg]) return false;
  r.s = seg_of(coff, nseg, c);
  r.c0 = coff[r.s];
  r.a = seg[r.s] + (c - r.c0) * SS_C;
  r.b = min(r.a + (uint32_t)SS_C, seg[r.s + 1]);
  return true;
}

// chunk sums of the terms in fp64 (csum[c]); with csq, mode 0 also stores the sums of their squares, from
// which the variance pass (mode 2) predicts its chunk sums as S2 - 2 m S1 + len m^2 without a pass of its own
__global__ __launch_bounds__(64) void ss_chunksum_k(const float *__restrict__ v, const uint32_t *__restrict__ seg,
                                                     const uint32_t *__restrict__ coff, int nseg, int mode,
                                                     const float *__restrict__ mean_dev, double *__restrict__ csum,
                                                     double *__restrict__ csq) {
  ChunkRef r;
  if (!chunk_ref(seg, coff, nseg, blockIdx.x, r)) return;
  const float mean = ss_mean(mode, mean_dev, seg, r.s);
  double acc = 0.0, acc2 = 0.0;
#pragma unroll
  for (int q = 0; q < SS_LANE; ++q) {
    const uint32_t i = r.a + (uint32_t)lane() * SS_LANE + q;
    if (i < r.b) {
      const double t = (double)ss_term(v[i], mode, mean);
      acc += t;
      acc2 += t * t;
    }
  }
  acc = wave_sum_f64(acc);
  if (csq) acc2 = wave_sum_f64(acc2);
  if (lane() == 0) {
    csum[blockIdx.x] = acc;
    if (csq) csq[blockIdx.x] = acc2;
  }
}

// predicted fp64 sum of chunk j's terms: csum[j], or in mode 2 from the moments of mode 0's terms
__device__ __forceinline__ double ss_chunk_pred(const double *__restrict__ csum, const double *__restrict__ csq,
                                                uint32_t j, double m, double len) {
  return csq ? csq[j] - 2.0 * m * csum[j] + len * m * m : csum[j];
}

__global__ __launch_bounds__(64) void ss_compose_k(const float *__restrict__ v, const uint32_t *__restrict__ seg,
                                                    const uint32_t *__restrict__ coff, int nseg, int mode,
                                                    const float *__restrict__ mean_dev,
                                                    const double *__restrict__ csum,
                                                    const double *__restrict__ csq, SSChunk *__restrict__ out,
                                                    int4 *__restrict__ lanes) {
  const uint32_t c = blockIdx.x;
  ChunkRef r;
  if (!chunk_ref(seg, coff, nseg, c, r)) return;
  const int l = lane();
  // the chunk's terms are read before the prefix sum, so the two round trips overlap
  const float mean = ss_mean(mode, mean_dev, seg, r.s);
  float x[SS_LANE];
#pragma unroll
  for (int q = 0; q < SS_LANE; ++q) {
    const uint32_t i = r.a + (uint32_t)l * SS_LANE + q;
    x[q] = i < r.b ? ss_term(v[i], mode, mean) : 0.0f;  // map (0, 0): the identity
  }
  // the segment's earlier chunk sums, four loads in flight per lane (chunks before c are full)
  const double md = (double)mean;
  double p = 0.0;
  for (uint32_t j0 = r.c0; j0 < c; j0 += 4 * WAVE) {
    double t[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t j = j0 + (uint32_t)(u * WAVE + l);
      t[u] = j < c ? ss_chunk_pred(csum, csq, j, md, (double)SS_C) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) p += t[u];
  }
  p = wave_sum_f64(p);
  const double lo = p, hi = p + ss_chunk_pred(csum, csq, c, md, (double)(r.b - r.a));
  int e = 0, kind = 0;  // 1: one binade (a chunk map); 2: two adjacent binades (lane maps for the crossing)
  if (c > r.c0 && lo > 0.0 && hi < 1e38) {
    const int elo = ilogb(lo * (1.0 - 1e-3)), ehi = ilogb(hi * (1.0 + 1e-3));
    if (elo == ehi && elo >= -125 && elo <= 126) {
      kind = 1;
      e = elo;
    } else if (ehi == elo + 1 && elo >= -125 && ehi <= 126 && lanes) {
      kind = 2;
      e = elo;
    }
  }
  if (kind == 0) {
    if (l == 0) out[c] = SSChunk{0, 0, 0, 0};
    return;
  }
  if (kind == 2) {
    // A chunk predicted to carry the sum from binade e into e + 1: per lane, the inclusive prefix of the lane
    // maps in binade e (lanes 0..l) and the composition of the later lanes' maps in binade e + 1 (lanes
    // l+1..63), so that the walk finds the crossing lane and finishes the chunk without building maps.
    int fe = 0, fo = 0, ge = 0, go = 0;
#pragma unroll
    for (int q = 0; q < SS_LANE; ++q) {
      int ae, ao;
      ss_term_map(x[q], 23 - e, ae, ao);
      ss_compose(fe, fo, ae, ao);
      ss_term_map(x[q], 22 - e, ae, ao);
      ss_compose(ge, go, ae, ao);
    }
    ss_scan_maps(fe, fo);
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {  // inclusive suffix: lanes l .. 63
      const int pe = __shfl_down(ge, d), po = __shfl_down(go, d);
      if (l + d < WAVE) ss_compose(ge, go, pe, po);
    }
    int se = __shfl_down(ge, 1), so = __shfl_down(go, 1);  // lanes l+1 .. 63
    if (l == WAVE - 1) se = so = 0;
    lanes[(size_t)c * WAVE + l] = make_int4(fe, fo, se, so);
    if (l == 0) out[c] = SSChunk{0, 0, e, 2};
    return;
  }
  int fe = 0, fo = 0;
#pragma unroll
  for (int q = 0; q < SS_LANE; ++q) {
    int ae, ao;
    ss_term_map(x[q], 23 - e, ae, ao);
    ss_compose(fe, fo, ae, ao);
  }
  // ordered tree: lane l composes its segment with the one of lane l + d
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const int ge = __shfl_down(fe, d), go = __shfl_down(fo, d);
    if ((l & (2 * d - 1)) == 0 && l + d < WAVE) ss_compose(fe, fo, ge, go);
  }
  if (l == 0) out[c] = SSChunk{fe, fo, e, 1};
}

// One chunk whose map does not apply, a lane's SS_LANE terms at a time.  While the running sum is a
// positive normal number, the maps of the remaining lanes' terms in its binade are built and scanned, and
// applied up to the first lane whose state leaves the binade (M >= 2^24); that lane alone is added term by
// term with the fp32 adder (the reference's own loop), which carries the sum into the next binade.  A +0
// sum skips the lanes whose terms are all zero (+0 plus +-0 is +0).  Every other state (subnormal,
// non-finite, negative) is also advanced with the fp32 adder.
struct SSTerms {
  float x[SS_LANE];
};

// the chunk's terms starting at element a, SS_LANE per lane; + 0 past the segment leaves the sum unchanged
__device__ __forceinline__ SSTerms ss_load_terms(const float *__restrict__ v, uint32_t a, uint32_t e_end, int mode,
                                                 float mean) {
  SSTerms t;
#pragma unroll
  for (int q = 0; q < SS_LANE; ++q) {
    const uint32_t i = a + (uint32_t)lane() * SS_LANE + q;
    t.x[q] = i < e_end ? ss_term(v[i], mode, mean) : 0.0f;
  }
  return t;
}

__device__ __forceinline__ float ss_replay(float acc, const SSTerms &t, int start = 0) {
  const int l = lane();
  const float *x = t.x;
  bool nz = false;
#pragma unroll
  for (int q = 0; q < SS_LANE; ++q) nz = nz || x[q] != 0.0f;
  int pos = start;
  while (pos < WAVE) {
    if (acc >= 0x1p-126f && acc < 0x1p127f) {
      const int e = ilogbf(acc);
      int fe = 0, fo = 0;  // identity for the lanes already added
      if (l >= pos) {
#pragma unroll
        for (int q = 0; q < SS_LANE; ++q) {
          int ae, ao;
          ss_term_map(x[q], 23 - e, ae, ao);
          ss_compose(fe, fo, ae, ao);
        }
      }
      ss_scan_maps(fe, fo);
      const int M0 = (int)ldexpf(acc, 23 - e);
      const int Ml = M0 + ((M0 & 1) ? fo : fe);
      const uint64_t bad = ballot(l >= pos && Ml >= (1 << 24));
      const int upto = bad ? __ffsll((unsigned long long)bad) - 1 : WAVE;  // lanes [pos, upto) accepted
      if (upto > pos) {
        acc = ldexpf((float)__builtin_amdgcn_readlane(Ml, upto - 1), e - 23);
        pos = upto;
      }
    } else if (__float_as_uint(acc) == 0u) {
      const uint64_t live = ballot(l >= pos && nz);
      pos = live ? __ffsll((unsigned long long)live) - 1 : WAVE;
    }
    if (pos >= WAVE) break;
#pragma unroll
    for (int r = 0; r < SS_LANE; ++r) acc = acc + readlanef(x[r], pos);
    ++pos;
  }
  return acc;
}

// A chunk predicted to cross from binade e into e + 1 (lane maps from ss_compose_k, m = this lane's (prefix
// map in e, map of the later lanes in e + 1)): the states after every lane in binade e in one step, the first
// lane whose state leaves the binade replayed term by term by the fp32 adder, and the later lanes as one map
// in binade e + 1, checked against 2^24 as every applied map is.  Anything else goes to ss_replay.
__device__ __forceinline__ float ss_cross(float acc, const SSTerms &t, int4 m, int e) {
  if (!(acc >= 0x1p-126f && acc < 0x1p127f) || ilogbf(acc) != e) return ss_replay(acc, t);
  const int M0 = (int)ldexpf(acc, 23 - e);
  const int Ml = M0 + ((M0 & 1) ? m.y : m.x);
  const uint64_t bad = ballot(Ml >= (1 << 24));
  if (!bad) return ldexpf((float)__builtin_amdgcn_readlane(Ml, WAVE - 1), e - 23);
  const int c = __ffsll((unsigned long long)bad) - 1;  // lanes [0, c) stay in binade e
  if (c > 0) acc = ldexpf((float)__builtin_amdgcn_readlane(Ml, c - 1), e - 23);
#pragma unroll
  for (int r = 0; r < SS_LANE; ++r) acc = acc + readlanef(t.x[r], c);
  if (c == WAVE - 1) return acc;
  if (acc >= 0x1p-126f && acc < 0x1p127f && ilogbf(acc) == e + 1) {
    const int M1 = (int)ldexpf(acc, 22 - e);
    const int se = __builtin_amdgcn_readlane(m.z, c), so = __builtin_amdgcn_readlane(m.w, c);
    const int M2 = M1 + ((M1 & 1) ? so : se);
    if (M2 < (1 << 24)) return ldexpf((float)M2, e - 22);
  }
  return ss_replay(acc, t, c + 1);
}

// One wave per segment walks its chunks in order, 64 at a time (lane j holds chunk j's map).  A run of
// consecutive clean chunks predicted in the running sum's binade is applied at once: an inclusive scan
// composes their maps across the lanes, every lane evaluates the state after its chunk from the run's
// start state, and the run is accepted up to the first chunk whose state leaves the binade (M >= 2^24;
// the sum never decreases, so every earlier state stayed inside).  That chunk, and any chunk outside such
// a run, goes through ss_replay.
__global__ __launch_bounds__(64) void ss_drive_k(const float *__restrict__ v, const uint32_t *__restrict__ seg,
                                                  const uint32_t *__restrict__ coff, int nseg, int mode,
                                                  const float *__restrict__ mean_dev,
                                                  const SSChunk *__restrict__ ch, const int4 *__restrict__ lanes,
                                                  float *__restrict__ out) {
  const int s = blockIdx.x;
  if (s >= nseg) return;
  const int l = lane();
  const float mean = ss_mean(mode, mean_dev, seg, s);
  const uint32_t c_begin = coff[s], c_end = coff[s + 1], e_end = seg[s + 1];
  float acc = 0.0f;
  const uint32_t e_begin = seg[s];
  // the next window's maps and the terms of the next chunk that is known to be replayed (one the binade
  // prediction left out) are loaded ahead, so their latency overlaps the work before them
  SSChunk mine = (c_begin + l < c_end) ? ch[c_begin + l] : SSChunk{0, 0, 0, 0};
  for (uint32_t c0 = c_begin; c0 < c_end; c0 += WAVE) {
    const SSChunk nxt = (c0 + WAVE + l < c_end) ? ch[c0 + WAVE + l] : SSChunk{0, 0, 0, 0};
    const int cnt = (int)min<uint32_t>(WAVE, c_end - c0);
    const uint64_t unclean = ballot(l < cnt && mine.clean != 1);
    int pf = unclean ? __ffsll((unsigned long long)unclean) - 1 : -1;
    SSTerms pft;
    int4 pfm = make_int4(0, 0, 0, 0);  // the crossing lane maps of chunk pf, when it has them
    if (pf >= 0) {
      pft = ss_load_terms(v, e_begin + (c0 + pf - c_begin) * SS_C, e_end, mode, mean);
      if (__builtin_amdgcn_readlane(mine.clean, pf) == 2) pfm = lanes[(size_t)(c0 + pf) * WAVE + l];
    }
    int pos = 0;
    while (pos < cnt) {
      // binade of the running sum (normal, finite, positive), else no map applies
      int e_acc = -100000;
      if (acc >= 0x1p-126f && acc < 0x1p127f) e_acc = ilogbf(acc);
      const uint64_t elig = ballot(l >= pos && l < cnt && mine.clean == 1 && mine.e == e_acc);
      const uint64_t stop = ~elig & (~0ull << pos);
      const int run_end = stop ? min(cnt, __ffsll((unsigned long long)stop) - 1) : cnt;
      if (run_end > pos) {
        int fe = 0, fo = 0;  // identity outside the run
        if (l >= pos && l < run_end) {
          fe = mine.ie;
          fo = mine.io;
        }
        ss_scan_maps(fe, fo);
        const int M0 = (int)ldexpf(acc, 23 - e_acc);
        const int Ml = M0 + ((M0 & 1) ? fo : fe);
        const uint64_t bad = ballot(l >= pos && l < run_end && Ml >= (1 << 24));
        const int upto = bad ? __ffsll((unsigned long long)bad) - 1 : run_end;  // chunks [pos, upto) accepted
        if (upto > pos) {
          acc = ldexpf((float)__builtin_amdgcn_readlane(Ml, upto - 1), e_acc - 23);
          pos = upto;
          continue;
        }
      }
      const int kind = __builtin_amdgcn_readlane(mine.clean, pos), ek = __builtin_amdgcn_readlane(mine.e, pos);
      if (pos == pf) {
        acc = kind == 2 ? ss_cross(acc, pft, pfm, ek) : ss_replay(acc, pft);
        const uint64_t later = unclean & (~0ull << pos) & ~(1ull << pos);
        pf = later ? __ffsll((unsigned long long)later) - 1 : -1;
        if (pf >= 0) {
          pft = ss_load_terms(v, e_begin + (c0 + pf - c_begin) * SS_C, e_end, mode, mean);
          if (__builtin_amdgcn_readlane(mine.clean, pf) == 2) pfm = lanes[(size_t)(c0 + pf) * WAVE + l];
        }
      } else {
        const SSTerms tt = ss_load_terms(v, e_begin + (c0 + pos - c_begin) * SS_C, e_end, mode, mean);
        acc = kind == 2 ? ss_cross(acc, tt, lanes[(size_t)(c0 + pos) * WAVE + l], ek) : ss_replay(acc, tt);
      }
      ++pos;
    }
    mine = nxt;
  }
  if (l == 0) out[s] = acc;
}

// chunk offsets of the segments (coff[nseg + 1], exclusive scan of ceil(len / SS_C)); returns the bound
static int ss_chunk_offsets(const uint32_t *seg_off, int nseg, size_t n_bound, Tmp &coff, size_t *bound,
                            hipStream_t s) {
  *bound = n_bound / SS_C + (size_t)nseg + 1;  // chunks: sum of ceil(len / SS_C)
  RVP_TRY(coff.alloc((nseg + 1) * sizeof(uint32_t), s));
  if (nseg <= 4096) {  // one wave walks the segments: one launch instead of two
    RVP_LAUNCH("ss_chunk_offsets_k", s, ss_chunk_offsets_k<<<1, 64, 0, s>>>(seg_off, nseg, coff.as<uint32_t>()));
    return RVP_OK;
  }
  RVP_LAUNCH("ss_nchunks_k", s, ss_nchunks_k<<<ceil_div(nseg, 256), 256, 0, s>>>(seg_off, nseg, coff.as<uint32_t>()));
  return exclusive_scan_u32(coff.as<uint32_t>(), coff.as<uint32_t>(), nseg, coff.as<uint32_t>() + nseg, s);
}

// One sum.  moments (2 * bound doubles) links a mode 0 run to the mode 2 run after it: mode 0 fills it
// with the chunk sums and sums of squares of its terms, mode 2 predicts from it and skips its own
// chunk-sum pass.
static int ss_run(const float *v, const uint32_t *seg_off, const uint32_t *coff, int nseg, size_t bound, int mode,
                  const float *mean_dev, float *out, double *moments, hipStream_t s) {
  Tmp cs, ch, lanes;
  double *csum = moments, *csq = moments ? moments + bound : nullptr;
  if (!moments) {
    RVP_TRY(cs.alloc(bound * sizeof(double), s));
    csum = cs.as<double>();
  }
  RVP_TRY(ch.alloc(bound * sizeof(SSChunk), s));
  RVP_TRY(lanes.alloc(bound * WAVE * sizeof(int4), s));
  if (mode != 2 || !moments)
    RVP_LAUNCH("ss_chunksum_k", s, ss_chunksum_k<<<(unsigned)bound, 64, 0, s>>>(v, seg_off, coff, nseg, mode,
                                                                                mean_dev, csum, csq));
  RVP_LAUNCH("ss_compose_k", s, ss_compose_k<<<(unsigned)bound, 64, 0, s>>>(v, seg_off, coff, nseg, mode, mean_dev,
                                                                            csum, mode == 2 ? csq : nullptr,
                                                                            ch.as<SSChunk>(), lanes.as<int4>()));
  RVP_LAUNCH("ss_drive_k", s, ss_drive_k<<<nseg, 64, 0, s>>>(v, seg_off, coff, nseg, mode, mean_dev,
                                                             ch.as<SSChunk>(), lanes.as<int4>(), out));
  return RVP_OK;
}

int exact_seq_sum_seg(const float *v, const uint32_t *seg_off, int nseg, size_t n_bound, int mode,
                      const float *mean_dev, float *out, hipStream_t s) {
  if (nseg <= 0) return RVP_OK;
  Tmp coff;
  size_t bound = 0;
  RVP_TRY(ss_chunk_offsets(seg_off, nseg, n_bound, coff, &bound, s));
  return ss_run(v, seg_off, coff.as<uint32_t>(), nseg, bound, mode, mean_dev, out, nullptr, s);
}

int exact_seq_mean_var_seg(const float *v, const uint32_t *seg_off, int nseg, size_t n_bound, float *sum_out,
                           float *sq_out, hipStream_t s) {
  if (nseg <= 0) return RVP_OK;
  Tmp coff, moments;
  size_t bound = 0;
  RVP_TRY(ss_chunk_offsets(seg_off, nseg, n_bound, coff, &bound, s));
  RVP_TRY(moments.alloc(2 * bound * sizeof(double), s));
  RVP_TRY(ss_run(v, seg_off, coff.as<uint32_t>(), nseg, bound, 0, nullptr, sum_out, moments.as<double>(), s));
  return ss_run(v, seg_off, coff.as<uint32_t>(), nseg, bound, 2, sum_out, sq_out, moments.as<double>(), s);
}

__global__ void one_segment_k(uint32_t n, uint32_t *__restrict__ seg) {
  if (threadIdx.x == 0) {
    seg[0] = 0;
    seg[1] = n;
  }
}

int exact_seq_sum(const float *v, size_t n, int mode, const float *mean_dev, float *out_dev, hipStream_t s) {
  RVP_CHECK(n < 0xffffff00ull, RVP_E_UNSUPPORTED, "sequential sum: more than 2^32 - 256 terms");
  Tmp seg;
  RVP_TRY(seg.alloc(2 * sizeof(uint32_t), s));
  RVP_LAUNCH("one_segment_k", s, one_segment_k<<<1, 64, 0, s>>>((uint32_t)n, seg.as<uint32_t>()));
  return exact_seq_sum_seg(v, seg.as<uint32_t>(), 1, n, mode, mean_dev, out_dev, s);
}

}  // namespace rvp

extern "C" int rvp_seq_sum(const float *v, size_t n, int squared_deviation, const float *mean_dev, float *out_dev,
                           rvp_stream_t st) {
  RVP_CHECK(out_dev && (v || n == 0) && (!squared_deviation || mean_dev), RVP_E_ARG, "null");
  RVP_TRY(rvp::pool_init());
  return rvp::exact_seq_sum(v, n, squared_deviation ? 1 : 0, mean_dev, out_dev, rvp::S(st));
}
