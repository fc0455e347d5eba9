// tune.cpp -- the SpMV autotuner.
//
// For one matrix and a list of block widths it chooses, by timing the product
// on the device:
//   * the SELL-C-sigma layout: chunk height C and sorting window sigma
//     (optional, rebuilds the device copy from the host CRS per candidate);
//   * per block width, the x bytes each lane gathers per nonzero, i.e. how
//     many lanes share a row and how many rows a wave covers (one scalar, or
//     one, two or four 16-byte vectors per lane);
//   * per block width, the workgroups per CU of the persistent grid (the
//     number of waves in flight, which trades latency hiding against the L2
//     footprint of the gathered x rows);
//   * per block width, for rows of at least 16 entries, the depth: 8, 16 or
//     32 entries of a row whose index loads and gathers go out together (16
//     and 32 also prefetch the next batch's indices);
//   * per block width above 16 columns, launches of 16 (and above 32, of 32)
//     columns instead of one: each reads the matrix again, but gathers rows
//     of x half (a quarter) as wide, so fewer of them leave L2;
//   * per block width, where every workgroup tile's x window fits LDS, the
//     product that stages each tile's window in LDS and gathers from there.
// The choice is stored on the matrix (SparseMat::tuned, used by every later
// SpMV of that width unless a global setting overrides it) and cached for
// the process by a matrix signature (scalar size, rows, entries, longest row,
// a bucket of the mean distance of entries from the diagonal, block width),
// so a matrix of the same signature is set up without timing again.
//
// Reference: PHIST chooses C and sigma statically when it builds a GHOST
// matrix (src/kernels/ghost/kernels.cpp:69-135); its builtin backend has no
// tunable layout.  Every candidate here yields the same product bit for bit
// (a row's sum keeps its entry order); dots fused into the product can
// differ in rounding between settings, as they differ between grid sizes.
#include "objects.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

namespace phx {

namespace {

struct Signature {
  int scalar;
  int64_t nrows, nnz;
  int32_t maxlen;
  int band;  // floor(log2(1 + mean |col - row|)) over sampled rows
  int nvec;
  bool operator<(const Signature& o) const {
    return std::tie(scalar, nrows, nnz, maxlen, band, nvec) <
           std::tie(o.scalar, o.nrows, o.nnz, o.maxlen, o.band, o.nvec);
  }
};
struct Choice {
  int C, sigma;
  SpmvTuned launch;
};
std::map<Signature, Choice>& cache() {
  static std::map<Signature, Choice> c;
  return c;
}
std::mutex& cache_mutex() {
  static std::mutex m;
  return m;
}

template <typename T>
int band_bucket(const SparseMat<T>* A) {
  const int64_t n = A->nrows;
  if (n == 0 || A->nnz == 0) return 0;
  const int64_t step = std::max<int64_t>(1, n / 4096);
  double sum = 0.0;
  int64_t cnt = 0;
  for (int64_t i = 0; i < n; i += step)
    for (int64_t p = A->crs_ptr[i]; p < A->crs_ptr[i + 1]; p++) {
      sum += std::fabs((double)A->crs_col[p] - (double)i);
      cnt++;
    }
  return cnt ? (int)std::floor(std::log2(1.0 + sum / (double)cnt)) : 0;
}

// scratch blocks for the trials (zero-filled: the timing does not depend on
// the values)
template <typename T>
struct TrialBlock {
  Mvec<T> v;
  size_t bytes = 0;
  TrialBlock(const Map* map, int64_t nrows, int nvec) {
    v.map = map;
    v.nrows = nrows;
    v.nvec = nvec;
    v.lda = nvec;
    bytes = sizeof(T) * (size_t)std::max<int64_t>(1, nrows) * nvec;
    v.dev_base = (T*)dev_alloc(bytes);
    PHX_HIP(hipMemsetAsync(v.dev_base, 0, bytes, rt().stream));
  }
  ~TrialBlock() { dev_release(v.dev_base, bytes); }
  TrialBlock(const TrialBlock&) = delete;
  TrialBlock& operator=(const TrialBlock&) = delete;
};

// fastest of reps products y = A x after one untimed launch; lb/bpc/depth < 0
// take the library's usual choice (global settings, tuned entry, table)
template <typename T>
double time_product(const SparseMat<T>* A, Mvec<T>* x, Mvec<T>* y, int lb, int bpc, int depth, int reps,
                    int col_block = -1, int stage = -1) {
  Runtime& r = rt();
  SpmvRequest rq;
  rq.lane_bytes = lb;
  rq.blocks_per_cu = bpc;
  rq.depth = depth;
  rq.col_block = col_block;
  rq.stage = stage;
  spmv_launch<T>(A, T(1), nullptr, x, T(0), y, rq, nullptr, nullptr);
  double best = 1e300;
  for (int k = 0; k < reps; k++) {
    PHX_HIP(hipEventRecord(r.ev_start, r.stream));
    spmv_launch<T>(A, T(1), nullptr, x, T(0), y, rq, nullptr, nullptr);
    PHX_HIP(hipEventRecord(r.ev_stop, r.stream));
    PHX_HIP(hipEventSynchronize(r.ev_stop));
    float ms = 0.f;
    PHX_HIP(hipEventElapsedTime(&ms, r.ev_start, r.ev_stop));
    best = std::min(best, (double)ms);
  }
  return best;
}

// the scalars per lane the SpMV takes for a block of nvec contiguous columns
// (lda = nvec, 16-byte aligned) and lane_bytes lb; candidates that give the
// same kernel are tried once
template <typename T>
int lanes_vpl(int nvec, int lb) {
  constexpr int V16 = 16 / sizeof(T);
  const bool vec_ok = nvec % V16 == 0;
  if (lb > 0 && lb <= (int)sizeof(T)) return 1;
  const int w = lb / (int)sizeof(T);
  if (vec_ok && (w == 2 * V16 || w == 4 * V16) && nvec % w == 0) return w;
  return vec_ok ? V16 : 1;
}

template <typename T>
SpmvTuned search_launch(const SparseMat<T>* A, int nvec, int reps) {
  TrialBlock<T> x(A->col_map, A->ncols, nvec), y(A->row_map, A->nrows, nvec);
  SpmvTuned best;
  best.ms = 1e300;
  std::vector<int> seen;
  std::vector<int> depths{8};
  if (A->max_row_len >= 16) depths.push_back(16);
  if (A->max_row_len >= 24) depths.push_back(32);
  std::vector<int> blocks{0};
  for (int w : {16, 32})
    if (nvec > w) blocks.push_back(w);
  for (int cbw : blocks) {
    seen.clear();
    const int ncol = cbw > 0 ? cbw : nvec;  // the widest launch
    for (int lb : {0, (int)sizeof(T), 32, 64}) {
      const int vpl = lanes_vpl<T>(nvec, lb);
      if (std::find(seen.begin(), seen.end(), vpl) != seen.end()) continue;
      seen.push_back(vpl);
      if (cbw > 0 && cbw % vpl != 0) continue;
      for (int stg : {0, 1}) {
        // staged windows where every tile's window fits LDS (depth 8, or 16 for lanes of at most 32 bytes)
        if (stg && !spmv_stage_for(A, ncol, vpl, nullptr, nullptr)) continue;
        std::vector<int> tried;
        for (int d : depths) {
          const int depth = stg ? (d >= 16 && vpl * (int)sizeof(T) <= 32 ? 16 : 8)
                                : spmv_effective_depth(d, vpl, (int)sizeof(T), ncol);
          if (std::find(tried.begin(), tried.end(), depth) != tried.end()) continue;
          tried.push_back(depth);
          for (int bpc : {1, 2, 3, 4, 6, 8}) {
            const double ms = time_product<T>(A, &x.v, &y.v, lb, bpc, depth, reps, cbw, stg);
            if (ms < best.ms) {
              best.ms = ms;
              best.lane_bytes = lb;
              best.blocks_per_cu = bpc;
              best.depth = depth;
              best.col_block = cbw;
              best.stage = stg;
            }
          }
        }
      }
    }
  }
  return best;
}

}  // namespace

// settings[7*i .. 7*i+6] = C, sigma, lane bytes, workgroups per CU, depth, column block, staging for nvecs[i]
template <typename T>
void sparse_tune(SparseMat<T>* A, int n_nvec, const int* nvecs, int flags, double* default_ms, double* tuned_ms,
                 int* settings) {
  ensure_init();
  PHX_REQUIRE(A, PHIST_INVALID_INPUT, "NULL matrix");
  PHX_REQUIRE(n_nvec >= 1 && n_nvec <= 16 && nvecs, PHIST_INVALID_INPUT, "%d block widths (1..16)", n_nvec);
  for (int i = 0; i < n_nvec; i++)
    PHX_REQUIRE(nvecs[i] >= 1 && nvecs[i] <= 256, PHIST_INVALID_INPUT, "block width %d (1..256)", nvecs[i]);
  PHX_REQUIRE(!rt().capturing, PHIST_FUNCTIONAL_ERROR, "the autotuner cannot run under graph capture");
  const bool layouts = flags & 1, fresh = flags & 2;
  const int reps = 3;

  // the library's untuned choice on the current layout, for the report
  A->tuned.clear();
  for (int i = 0; i < n_nvec; i++) {
    TrialBlock<T> x(A->col_map, A->ncols, nvecs[i]), y(A->row_map, A->nrows, nvecs[i]);
    const double ms = time_product<T>(A, &x.v, &y.v, -1, -1, -1, reps);
    if (default_ms) default_ms[i] = ms;
  }

  const int band = band_bucket(A);
  std::vector<Signature> sig(n_nvec);
  for (int i = 0; i < n_nvec; i++)
    sig[i] = Signature{(int)sizeof(T), A->nrows, A->nnz, A->max_row_len, band, nvecs[i]};

  // cached: every width known under one layout
  std::vector<Choice> chosen(n_nvec);
  bool hit = !fresh;
  if (hit) {
    std::lock_guard<std::mutex> lk(cache_mutex());
    for (int i = 0; i < n_nvec && hit; i++) {
      auto it = cache().find(sig[i]);
      hit = it != cache().end() && (i == 0 || (it->second.C == chosen[0].C && it->second.sigma == chosen[0].sigma));
      if (hit) chosen[i] = it->second;
    }
  }
  if (!hit) {
    std::vector<std::pair<int, int>> cands{{A->C, A->sigma}};
    if (layouts) {
      bool uniform = true;
      for (int64_t i = 1; i < A->nrows && uniform; i++)
        uniform = A->crs_ptr[i + 1] - A->crs_ptr[i] == A->crs_ptr[1] - A->crs_ptr[0];
      for (int C : {16, 32, 64})
        for (int sg : {1, 256}) {
          if (uniform && sg != 1) continue;  // sorting equal rows changes nothing
          std::pair<int, int> c{C, sg};
          if (std::find(cands.begin(), cands.end(), c) == cands.end()) cands.push_back(c);
        }
    }
    double best_total = 1e300;
    for (const auto& c : cands) {
      if (c.first != A->C || c.second != A->sigma) sell_rebuild(A, c.first, c.second);
      std::vector<Choice> here(n_nvec);
      double total = 0.0;
      for (int i = 0; i < n_nvec; i++) {
        here[i] = Choice{A->C, A->sigma, search_launch<T>(A, nvecs[i], reps)};
        total += here[i].launch.ms;
      }
      if (total < best_total) {
        best_total = total;
        chosen = here;
      }
    }
    std::lock_guard<std::mutex> lk(cache_mutex());
    for (int i = 0; i < n_nvec; i++) cache()[sig[i]] = chosen[i];
  }

  if (chosen[0].C != A->C || chosen[0].sigma != A->sigma) sell_rebuild(A, chosen[0].C, chosen[0].sigma);
  for (int i = 0; i < n_nvec; i++) A->tuned[nvecs[i]] = chosen[i].launch;
  for (int i = 0; i < n_nvec; i++) {
    TrialBlock<T> x(A->col_map, A->ncols, nvecs[i]), y(A->row_map, A->nrows, nvecs[i]);
    const double ms = time_product<T>(A, &x.v, &y.v, -1, -1, -1, reps);
    A->tuned[nvecs[i]].ms = ms;
    if (tuned_ms) tuned_ms[i] = ms;
    if (settings) {
      settings[7 * i + 0] = A->C;
      settings[7 * i + 1] = A->sigma;
      settings[7 * i + 2] = chosen[i].launch.lane_bytes;
      settings[7 * i + 3] = chosen[i].launch.blocks_per_cu;
      settings[7 * i + 4] = chosen[i].launch.depth;
      settings[7 * i + 5] = chosen[i].launch.col_block;
      settings[7 * i + 6] = chosen[i].launch.stage;
    }
  }
  sync_stream();
}

template void sparse_tune<double>(SparseMat<double>*, int, const int*, int, double*, double*, int*);
template void sparse_tune<float>(SparseMat<float>*, int, const int*, int, double*, double*, int*);

}  // namespace phx

using namespace phx;

extern "C" {

void phist_Damd_sparseMat_tune(phist_DsparseMat_ptr A, int n_nvec, const int* nvecs, int flags, double* default_ms,
                               double* tuned_ms, int* settings, int* iflag) {
  guard(iflag, "phist_Damd_sparseMat_tune", [&] {
    SparseMat<double>* a = SP<double>(A);
    sparse_tune<double>(a, n_nvec, nvecs, flags & 3, default_ms, tuned_ms, settings);
    if (flags & 4) {  // the FP32 twin of the mixed-precision solvers, same widths, launch settings only
      SparseMat<float>* f = const_cast<SparseMat<float>*>(sparse_fp32_twin(a));
      sparse_tune<float>(f, n_nvec, nvecs, flags & 2, nullptr, nullptr, nullptr);
    }
    *iflag = 0;
  });
}

void phist_Samd_sparseMat_tune(phist_SsparseMat_ptr A, int n_nvec, const int* nvecs, int flags, double* default_ms,
                               double* tuned_ms, int* settings, int* iflag) {
  guard(iflag, "phist_Samd_sparseMat_tune", [&] {
    sparse_tune<float>(SP<float>(A), n_nvec, nvecs, flags & 3, default_ms, tuned_ms, settings);
    *iflag = 0;
  });
}

}  // extern "C"
