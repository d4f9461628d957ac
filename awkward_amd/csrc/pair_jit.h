// pair_jit.h -- the pair-extrema kernel for one program, compiled at run time
// (hipRTC) by jit.hip: ak.max / ak.min / ak.argmax / ak.argmin at axis=-1 of
// an elementwise expression of ak.combinations(array, 2)
// (listoffsetarray.py:1152-1261, array_module.py:254-332,
// awkward_reduce_{max,min,argmax,argmin}.cpp).
//
// The including source defines, in akhip::pjit, the program's shape as
// constants -- W (register word), F (output float), kMX / kMN (which
// extrema), kStage (elements per staged column and wave), kU (pairs a lane
// evaluates together), kNLeaves, kNSlots, kSlot[] / kLeg[] / kType[] per
// leaf, kFirst[] per staged column -- and `eval(l)`, the program as
// straight-line code over the leaves' canonical values: each op is expr.h's
// x_op_v with the op as a compile-time constant, so every result is the
// interpreter's, bit for bit (the source is compiled with contraction off).
//
// One wave takes 64 lists, stages the distinct columns of as many of them as
// fit (kStage elements), spreads their pairs evenly over its lanes in
// combinations order (i < j by position, lexicographic), and each lane
// reduces its run of pairs in registers as it evaluates them.  "The first
// pair holding the largest non-NaN value" composes over consecutive runs of
// a list's pairs, and the CPU loops (awkward_reduce_max.cpp: from -inf, take
// x > best; awkward_reduce_argmax.cpp: seed with the first value, take
// x > best) give exactly that, except that argmax stays 0 when the first pair
// is NaN, since nothing compares greater than NaN.  A list that starts and
// ends inside one run is finished there; the first and last list of every
// run leave a partial that the list's own lane merges in run order.  A list
// longer than the stage is spread over the wave alone, its leaves read from
// memory.
#pragma once

namespace akhip {
namespace pjit {

struct Leaves {
  const void* data[kExprMaxLeaves];
  const int64_t* index[kExprMaxLeaves];
};

struct Out {
  void* max;  // F
  void* min;
  int64_t* argmax;  // local pair index, -1 for a list without pairs
  int64_t* argmin;
  int8_t* mask;  // 1 for a list without pairs
};

template <typename I>
struct Part {
  F vmax, vmin;
  I imax, imin;  // first pair of vmax / vmin among the non-NaN values; -1: none yet
  int nan0;      // the list's first pair is NaN (known to the lane that holds it)
};

template <typename I>
__device__ inline Part<I> part_empty() {
  return Part<I>{(F)-__builtin_inf(), (F)__builtin_inf(), (I)-1, (I)-1, 0};
}

template <typename I>
__device__ inline void part_take(Part<I>& a, F x, I r) {
  if constexpr (kMX) {
    if (a.imax < 0 ? x == x : x > a.vmax) {
      a.vmax = x;
      a.imax = r;
    }
  }
  if constexpr (kMN) {
    if (a.imin < 0 ? x == x : x < a.vmin) {
      a.vmin = x;
      a.imin = r;
    }
  }
  if (r == 0) a.nan0 = x != x;
}

// a, then the later pairs b: ties keep a
template <typename I>
__device__ inline void part_merge(Part<I>& a, const Part<I>& b) {
  if constexpr (kMX) {
    if (b.imax >= 0 && (a.imax < 0 || b.vmax > a.vmax)) {
      a.vmax = b.vmax;
      a.imax = b.imax;
    }
  }
  if constexpr (kMN) {
    if (b.imin >= 0 && (a.imin < 0 || b.vmin < a.vmin)) {
      a.vmin = b.vmin;
      a.imin = b.imin;
    }
  }
  a.nan0 |= b.nan0;
}

template <typename I>
__device__ inline Part<I> part_shfl(const Part<I>& a, int src) {
  Part<I> r;
  r.vmax = __shfl(a.vmax, src, kWave);
  r.vmin = __shfl(a.vmin, src, kWave);
  r.imax = __shfl(a.imax, src, kWave);
  r.imin = __shfl(a.imin, src, kWave);
  r.nan0 = __shfl(a.nan0, src, kWave);
  return r;
}

// The reducers' results for list b: for a list with pairs, max / min are
// -inf / +inf when every value is NaN (the identities, never replaced) and
// argmax / argmin are 0 when the first value is NaN; without pairs, the
// identities, -1 and mask 1.
template <typename I>
__device__ inline void part_store(const Out& out, int64_t b, const Part<I>& a, bool any) {
  if constexpr (kMX) {
    if (out.max) ((F*)out.max)[b] = a.imax < 0 ? (F)-__builtin_inf() : a.vmax;
    if (out.argmax) out.argmax[b] = !any ? -1 : a.nan0 ? 0 : (int64_t)a.imax;
  }
  if constexpr (kMN) {
    if (out.min) ((F*)out.min)[b] = a.imin < 0 ? (F)__builtin_inf() : a.vmin;
    if (out.argmin) out.argmin[b] = !any ? -1 : a.nan0 ? 0 : (int64_t)a.imin;
  }
  if (out.mask) out.mask[b] = any ? 0 : 1;
}

__device__ inline int wave_scan(int v) {
  const int lane = threadIdx.x;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int u = __shfl_up(v, d, kWave);
    if (lane >= d) v += u;
  }
  return v;
}

// Pair r (lexicographic) of a list of m items: i is the largest with
// i (2m - i - 1) / 2 <= r (the pairs before first element i), estimated from
// the quadratic and corrected with exact integer counts.
template <typename I>
__device__ inline void pair_at(I r, I m, I& i, I& j) {
  const double t = 2.0 * (double)m - 1.0;
  double disc = t * t - 8.0 * (double)r;
  disc = disc < 0.0 ? 0.0 : disc;
  I a = (I)((t - __builtin_sqrt(disc)) * 0.5);
  a = a < 0 ? 0 : (a > m - 2 ? m - 2 : a);
  while (a > 0 && a * (2 * m - a - 1) / 2 > r) a--;
  while (a < m - 2 && (a + 1) * (2 * m - a - 2) / 2 <= r) a++;
  i = a;
  j = a + 1 + (r - a * (2 * m - a - 1) / 2);
}

__device__ inline ExprLeaf leaf_of(const Leaves& lv, int k) { return ExprLeaf{lv.data[k], lv.index[k], kType[k], 0}; }

// One lane's run [q, qe) of the pairs of lists [b0, b1) (pair prefix pb[],
// element offsets ol[] from the stage's first element Eb), reduced as it is
// evaluated: lists finished inside the run go to res[]; the run's first list
// (he, hp) and last list (te, tp) come back as partials.  STAGED: the leaves
// come from the stage and the run may cross lists; else the run lies in one
// list (b1 = b0 + 1) whose leaves come from memory.
template <typename I, bool STAGED>
__device__ inline void run(const Leaves& lv, const W* s_col, int64_t Eb, const I* pb, const I* ol, int b0, int b1,
                           I q, I qe, Part<I>* res, int& he, Part<I>& hp, int& te, Part<I>& tp) {
  int lo = b0, hi = b1;  // the list holding pair q: the last e with pb[e] <= q
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pb[mid] <= q) lo = mid;
    else hi = mid;
  }
  int e = lo;
  I m = ol[e + 1] - ol[e], a = ol[e], i, j;
  pair_at<I>(q - pb[e], m, i, j);
  he = e;
  int cur = e;
  I rk = q - pb[e], cc = pb[e + 1] - pb[e], qr = q;
  bool first = true;
  Part<I> part = part_empty<I>();
  while (q < qe) {
    I pi[kU], pj[kU];
    bool on[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      on[u] = q < qe;
      pi[u] = a + i;
      pj[u] = a + j;
      if (on[u]) {
        q++;
        if (++j == m) {
          ++i;
          j = i + 1;
        }
        if constexpr (STAGED) {
          if (i >= m - 1 && q < qe) {  // on to the next list with pairs
            do {
              e++;
            } while (pb[e + 1] == pb[e]);
            m = ol[e + 1] - ol[e];
            a = ol[e];
            i = 0;
            j = 1;
          }
        }
      }
    }
    W v[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      W l[kNLeaves > 0 ? kNLeaves : 1];
#pragma unroll
      for (int k = 0; k < kNLeaves; k++) {
        const I at = kLeg[k] ? pj[u] : pi[u];
        if constexpr (STAGED) l[k] = on[u] ? s_col[kSlot[k] * kStage + (int)at] : W(0);
        else l[k] = on[u] ? x_load<W>(leaf_of(lv, k), Eb + at) : W(0);
      }
      v[u] = eval(l);
    }
#pragma unroll
    for (int u = 0; u < kU; u++) {
      if (qr < qe) {
        F x;
        if constexpr (sizeof(F) == 4) x = __uint_as_float((uint32_t)v[u]);
        else x = __longlong_as_double((long long)v[u]);
        part_take<I>(part, x, rk);
        qr++;
        rk++;
        if constexpr (STAGED) {
          if (rk == cc && qr < qe) {  // list cur ends inside the run
            if (first) hp = part;
            else res[cur] = part;
            first = false;
            do {
              cur++;
            } while (pb[cur + 1] == pb[cur]);
            rk = 0;
            cc = pb[cur + 1] - pb[cur];
            part = part_empty<I>();
          }
        }
      }
    }
  }
  te = cur;
  tp = part;
  if (first) hp = part;
}

// Lane e finishes list e of [b0, b1) in res[] unless the run holding its
// first pair did: the merge of the partials of the runs (lanes la .. lb) its
// pairs fall in.  Every lane runs the same number of merge steps, so the
// shuffles read active lanes.  Then lane e stores list e.  res[e - r0] holds
// list e.
template <typename I>
__device__ inline void finish(const I* pb, int b0, int b1, I per, int64_t L0, const Out& out, Part<I>* res, int r0,
                              int he, const Part<I>& hp, int te, const Part<I>& tp) {
  const int e = threadIdx.x;
  const bool mine = e >= b0 && e < b1;
  int la = 0, span = 0;
  if (mine) {
    const I c0 = pb[e], c1 = pb[e + 1];
    if (c1 > c0) {
      la = (int)(c0 / per);
      span = (int)((c1 - 1) / per) - la + 1;
    }
  }
  int steps = span;
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const int o = __shfl_xor(steps, d, kWave);
    steps = o > steps ? o : steps;
  }
  bool merged = span > 0;
  Part<I> acc = part_empty<I>();
  for (int k = 0; k < steps; k++) {
    const int src = la + k < kWave ? la + k : kWave - 1;
    const int she = __shfl(he, src, kWave), ste = __shfl(te, src, kWave);
    const Part<I> h = part_shfl<I>(hp, src), t = part_shfl<I>(tp, src);
    if (k < span) {
      if (k == 0 && e != she && e != ste) merged = false;  // finished inside run la
      part_merge<I>(acc, e == ste ? t : h);
    }
  }
  if (merged) res[e - r0] = acc;
  __syncthreads();
  if (mine) part_store<I>(out, L0 + e, span > 0 ? res[e - r0] : part_empty<I>(), span > 0);
}

}  // namespace pjit
}  // namespace akhip

extern "C" __global__ __launch_bounds__(64) void akhip_pair_jit(akhip::pjit::Leaves lv,
                                                                const int64_t* __restrict__ offsets, int64_t n,
                                                                akhip::pjit::Out out) {
  using namespace akhip;
  using namespace akhip::pjit;
  extern __shared__ uint64_t s_dyn[];
  W* s_col = (W*)s_dyn;  // kNSlots columns of kStage
  __shared__ int64_t s_meta[2 * (kWave + 1)];
  __shared__ Part<int> s_res[kWave];
  const int lane = threadIdx.x;
  const int64_t L0 = (int64_t)blockIdx.x * kWave;
  const int nl = (int)(n - L0 < kWave ? n - L0 : kWave);
  const int64_t olo = lane < nl ? offsets[L0 + lane] : 0;
  const int64_t ohi = lane < nl ? offsets[L0 + lane + 1] : 0;
  int b0 = 0;
  while (b0 < nl) {  // uniform: sub-batches of lists whose elements fit the stage
    const int64_t Eb = __shfl(olo, b0, kWave);
    const bool fits = lane >= b0 && lane < nl && olo >= Eb && ohi >= olo && ohi - Eb <= kStage;
    const uint64_t nofit = ~(uint64_t)__ballot(fits) & (~uint64_t(0) << b0);
    int b1 = nofit ? (int)__builtin_ctzll(nofit) : kWave;
    b1 = b1 < nl ? b1 : nl;
    int he = -1, te = -1;
    if (b1 > b0) {
      int* pb = (int*)s_meta;
      int* ol = pb + kWave + 1;
      const int span = (int)(__shfl(ohi, b1 - 1, kWave) - Eb);
      for (int q = lane; q < span; q += kWave) {
#pragma unroll
        for (int s = 0; s < kNSlots; s++) s_col[s * kStage + q] = x_load<W>(leaf_of(lv, kFirst[s]), Eb + q);
      }
      const bool in = lane >= b0 && lane < b1;
      const int m = in ? (int)(ohi - olo) : 0;
      const int c = m * (m - 1) / 2;
      const int incl = wave_scan(c);
      pb[lane] = incl - c;
      ol[lane] = lane < b0 ? 0 : (in ? (int)(olo - Eb) : span);
      if (lane == kWave - 1) {
        pb[kWave] = incl;
        ol[kWave] = span;
      }
      __syncthreads();
      const int T = pb[kWave];
      const int per = T > 0 ? (T + kWave - 1) / kWave : 1;
      const int qs = lane * per, qe = qs + per < T ? qs + per : T;
      Part<int> hp = part_empty<int>(), tp = hp;
      if (qs < qe) run<int, true>(lv, s_col, Eb, pb, ol, b0, b1, qs, qe, s_res, he, hp, te, tp);
      finish<int>(pb, b0, b1, per, L0, out, s_res, 0, he, hp, te, tp);
    } else {  // list b0 alone: longer than the stage
      b1 = b0 + 1;
      int64_t* pb = s_meta;
      int64_t* ol = s_meta + kWave + 1;
      const int64_t mm = __shfl(ohi, b0, kWave) - Eb;
      const int64_t m = mm > 0 ? mm : 0;
      const int64_t c = m * (m - 1) / 2;
      pb[lane] = lane <= b0 ? 0 : c;
      ol[lane] = lane <= b0 ? 0 : m;
      if (lane == kWave - 1) {
        pb[kWave] = c;
        ol[kWave] = m;
      }
      __syncthreads();
      const int64_t per = c > 0 ? (c + kWave - 1) / kWave : 1;
      const int64_t qs = lane * per, qe = qs + per < c ? qs + per : c;
      Part<int64_t> hp = part_empty<int64_t>(), tp = hp;
      if (qs < qe) run<int64_t, false>(lv, nullptr, Eb, pb, ol, b0, b1, qs, qe, nullptr, he, hp, te, tp);
      // the stage is free: one entry, list b0
      finish<int64_t>(pb, b0, b1, per, L0, out, (Part<int64_t>*)s_dyn, b0, he, hp, te, tp);
    }
    __syncthreads();  // the stage and s_meta are rewritten by the next sub-batch
    b0 = b1;
  }
}
