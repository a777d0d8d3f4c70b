// C ABI (include/warp_psr_amd.h): parameter checks with the reference's
// messages (src/warp_psr/reconstruction.py:181-234), stage orchestration with
// HIP-event stage timings (reconstruction.py:241-250), result ownership and
// the named stage-array export used by parity tests.
#include "../../include/warp_psr_amd.h"
#include "../../include/warp_psr_amd_field.h"
#include "psr_result.h"
#include "psr_mesh.h"
#include "psr_meta.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <map>
#include <sstream>
#include <atomic>
#include <chrono>
#include <thread>

using namespace psr;

struct PsrResult {
  int dtype = 0;
  void* impl = nullptr;
  std::map<std::string, std::vector<char>> cache;
  DArray<int> level_offsets_dev;
  std::string metadata;
  // built on a library worker stream (batched): its blocks return to that
  // stream's cache, so freeing waits for every stream that may still read it
  bool sync_on_free = false;
};

namespace {

thread_local std::string g_error;

int fail(const PsrError& e) {
  g_error = e.what();
  return e.code;
}

void check_params(const PsrParams& p) {
  const int depth = p.depth, full = p.full_depth, base = p.base_depth, solve = p.solve_depth, kd = p.kernel_depth;
  if (!(0 <= base && base <= full && full <= solve && solve <= depth && depth <= 15) || depth < 1 || !(0 <= kd && kd <= depth))
    throw PsrError(2, "Require 0 <= base_depth <= full_depth <= solve_depth <= depth <= 15, depth >= 1, and 0 <= kernel_depth <= depth");
  if (depth - full > 7)
    throw PsrError(2, "Warp 1.19 AdaptiveNanogrid supports at most 8 integration levels; require depth - full_depth <= 7");
  if (p.iters < 1 || p.base_v_cycles < 1) throw PsrError(2, "iters and base_v_cycles must be positive integers");
  if (!std::isfinite(p.point_weight) || p.point_weight < 0 || !std::isfinite(p.cg_accuracy) || p.cg_accuracy <= 0)
    throw PsrError(2, "Require finite nonnegative point_weight and positive cg_accuracy");
  if (!std::isfinite(p.samples_per_node) || p.samples_per_node <= 0 || !std::isfinite(p.scale) || p.scale <= 1)
    throw PsrError(2, "Require finite positive samples_per_node and scale > 1");
  if (!(p.relax_tile == 0 || p.relax_tile == -1 || (p.relax_tile >= 8 && p.relax_tile <= 40 && p.relax_tile % 2 == 0)))
    throw PsrError(2, "relax_tile must be 0 (default), -1 (one launch per colour) or an even tile side in [8, 40]");
}

Params resolve(const PsrParams& in) {
  PsrParams p = in;
  if (p.base_depth < 0) p.base_depth = p.full_depth;
  if (p.solve_depth < 0) p.solve_depth = p.depth;
  if (p.kernel_depth < 0) p.kernel_depth = std::max(p.depth - 2, 0);
  check_params(p);
  Params q;
  q.depth = p.depth; q.full_depth = p.full_depth; q.base_depth = p.base_depth; q.solve_depth = p.solve_depth;
  q.kernel_depth = p.kernel_depth; q.samples_per_node = p.samples_per_node; q.scale = p.scale;
  q.point_weight = p.point_weight; q.exact_interpolation = p.exact_interpolation ? 1 : 0; q.iters = p.iters;
  q.base_v_cycles = p.base_v_cycles; q.cg_accuracy = p.cg_accuracy; q.keep_intermediates = p.keep_intermediates;
  q.float_operators = p.float_operators ? 1 : 0;
  q.relax_tile = p.relax_tile;
  return q;
}

template <class T> Result<T>* run(const T* pts, const T* nrm, int64_t n, const Params& prm, hipStream_t s) {
  auto r = std::make_unique<Result<T>>();
  r->prm = prm;
  r->s = s;
  r->n = n;
  hipEvent_t ev[7];
  for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
  const char* names[6] = {"samples", "density", "normal_field", "tree", "assembly_and_solve", "field_and_isovalue"};
  try {
    HIP_CHECK(hipEventRecord(ev[0], s));
    stage_samples(*r, pts, nrm);
    HIP_CHECK(hipEventRecord(ev[1], s));
    stage_density(*r);
    HIP_CHECK(hipEventRecord(ev[2], s));
    stage_normals(*r);
    HIP_CHECK(hipEventRecord(ev[3], s));
    stage_tree(*r);
    HIP_CHECK(hipEventRecord(ev[4], s));
    stage_solve(*r);
    HIP_CHECK(hipEventRecord(ev[5], s));
    stage_field(*r);
    HIP_CHECK(hipEventRecord(ev[6], s));
    HIP_CHECK(hipEventSynchronize(ev[6]));
    for (int k = 0; k < 6; ++k) {
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
      r->timings.push_back({names[k], ms / 1000.f});
    }
    for (size_t k = 0; k < r->relax_stats.size(); ++k) {
      auto& st = r->relax_stats[k];
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, st.start, st.end));
      st.seconds = ms / 1000.f;
      unsigned long long nnz = 0;
      if (k < r->relax_nnz.size()) read_bytes(&nnz, r->relax_nnz[k].data(), sizeof(nnz), s);
      st.nnz = (int64_t)nnz;
    }
    // the base V-cycles whose root conjugate gradient ran (flags written on
    // the device, read here where the run has synchronised anyway)
    if (r->cg_ran.size()) {
      std::vector<int> ran(r->cg_ran.size());
      read_bytes(ran.data(), r->cg_ran.data(), sizeof(int) * ran.size(), s);
      for (size_t c = 0; c < ran.size(); ++c)
        if (ran[c]) r->cg_runs.push_back((int)c);
    }
  } catch (...) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    throw;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return r.release();
}

template <class X> void put(std::vector<char>& buf, const X* p, size_t n) {
  buf.resize(n * sizeof(X));
  if (n) std::memcpy(buf.data(), p, n * sizeof(X));
}
template <class X> void put_dev(std::vector<char>& buf, const DArray<X>& a) {
  std::vector<X> h = a.to_host();
  put(buf, h.data(), h.size());
}

// position of canonical row i in a schedule-ordered array (identity if none)
std::vector<int> inverse_order(const DArray<int>& order, int R) {
  std::vector<int> inv(R);
  if (order.empty()) {
    for (int i = 0; i < R; ++i) inv[i] = i;
    return inv;
  }
  std::vector<int> o = order.to_host();
  for (int j = 0; j < R; ++j) inv[o[j]] = j;
  return inv;
}

// stencil -> compact CSR in canonical row order (present slots in slot order)
// (values exported as Out: float operators of mixed runs widen to double)
template <class Out, class V>
void stencil_csr(const Stencil<V>& m, int slots, const char* part, std::vector<char>& out, const std::vector<int>& inv) {
  std::vector<int> cols = m.cols.to_host();
  std::vector<V> vals = m.vals.to_host();
  const int R = m.rows;
  std::vector<int> offs(R + 1, 0), cc;
  std::vector<Out> vv;
  for (int i = 0; i < R; ++i) {
    const int j = inv.empty() ? i : inv[i];
    for (int k = 0; k < slots; ++k) {
      int c = cols[(size_t)k * R + j];
      if (c >= 0) { cc.push_back(c); vv.push_back(Out(vals[(size_t)k * R + j])); }
    }
    offs[i + 1] = (int)cc.size();
  }
  if (!std::strcmp(part, "offsets")) put(out, offs.data(), offs.size());
  else if (!std::strcmp(part, "columns")) put(out, cc.data(), cc.size());
  else put(out, vv.data(), vv.size());
}

template <class T> bool export_array(Result<T>& r, const std::string& name, int level, std::vector<char>& out) {
  const bool lvl = level >= 0 && level < (int)r.levels.size();
  auto L = [&]() -> LevelState<T>& { return r.levels[level]; };
  if (name == "transform") { put(out, r.transform, 5); return true; }
  if (name == "bounds") { put(out, r.bounds, 6); return true; }
  if (name == "positions") { put_dev(out, r.q); return true; }
  if (name == "unit_normals") { put_dev(out, r.unit_normals); return true; }
  if (name == "input_status") { put_dev(out, r.input_status); return true; }
  if (name == "input_group_indices") { put_dev(out, r.input_group); return true; }
  if (name == "input_order") { put_dev(out, r.input_order); return true; }
  if (name == "sample_keys") { put_dev(out, r.sample_keys); return true; }
  if (name == "position_sums") { put_dev(out, r.position_sums); return true; }
  if (name == "normal_sums") { put_dev(out, r.normal_sums); return true; }
  if (name == "weights") { put_dev(out, r.weights); return true; }
  if (name == "first_indices") { put_dev(out, r.first_indices); return true; }
  if (name == "group_offsets") { put_dev(out, r.group_offsets); return true; }
  if (name == "aggregate_keys" && lvl) { put_dev(out, L().agg_keys); return true; }
  if (name == "aggregate_values" && lvl) { put_dev(out, L().agg_vals); return true; }
  if (name == "density_keys" && level >= 0 && level <= r.prm.kernel_depth) { put_dev(out, L().dens_keys); return true; }
  if (name == "density_values" && level >= 0 && level <= r.prm.kernel_depth) { put_dev(out, L().dens_vals); return true; }
  if (name == "densities") { put_dev(out, r.densities); return true; }
  if (name == "depths_and_weights") { put_dev(out, r.dw); return true; }
  if (name == "density_scale") { put(out, &r.density_scale, 1); return true; }
  if (name == "normal_keys") { put_dev(out, r.normal_keys); return true; }
  if (name == "normal_values") { put_dev(out, r.normal_vals); return true; }
  if (name == "normal_average") { put(out, r.normal_average_host, 3); return true; }
  if (name == "normal_offsets") { put(out, r.normal_offsets.data(), r.normal_offsets.size()); return true; }
  if (name == "tree_keys" && lvl) { put_dev(out, L().tree_keys); return true; }
  if (name == "basis_keys" && lvl) { put_dev(out, L().basis_keys); return true; }
  auto canonical = [&](const auto& a) {  // schedule-ordered vector -> canonical rows
    auto h = a.to_host();
    if (!L().scheduled) { put(out, h.data(), h.size()); return; }
    std::vector<int> o = L().order.to_host();
    decltype(h) c(h.size());
    for (size_t j = 0; j < o.size(); ++j) c[o[j]] = h[j];
    put(out, c.data(), c.size());
  };
  if (name == "normal_load" && lvl) { put_dev(out, L().normal_load); return true; }
  if (name == "point_load" && lvl) { canonical(L().point_load); return true; }
  if (name == "rhs" && lvl) { canonical(L().rhs); return true; }
  if (name == "corrections" && lvl) { put_dev(out, L().correction); return true; }
  const bool mat = lvl && (name.rfind("matrix_", 0) == 0 || name.rfind("cross_", 0) == 0);
  const std::vector<int> inv = mat && L().scheduled ? inverse_order(L().order, L().basis_count) : std::vector<int>();
  if (lvl && name.rfind("matrix_", 0) == 0) {
    if (L().A32.rows) stencil_csr<T>(L().A32, 27, name.c_str() + 7, out, inv);
    else stencil_csr<T>(L().A, 27, name.c_str() + 7, out, inv);
    return true;
  }
  if (lvl && name.rfind("cross_", 0) == 0) {  // canonical rows
    if (L().cross32.rows) stencil_csr<T>(L().cross32, 27, name.c_str() + 6, out, {});
    else stencil_csr<T>(L().cross, 27, name.c_str() + 6, out, {});
    return true;
  }
  if (lvl && name.rfind("prolong_", 0) == 0) { stencil_csr<T>(L().P, 8, name.c_str() + 8, out, {}); return true; }
  if (name == "keys") { put_dev(out, r.keys); return true; }
  if (name == "coefficients") { put_dev(out, r.coefficients); return true; }
  if (name == "cumulative") { put_dev(out, r.cumulative); return true; }
  if (name == "centroid_values") { put_dev(out, r.centroid_values); return true; }
  if (name == "iso") { put(out, &r.iso, 1); return true; }
  if (name == "cube_to_model") { put(out, r.cube_to_model, 16); return true; }
  if (name == "model_to_cube") { put(out, r.model_to_cube, 16); return true; }
  if (name == "cg_runs") { put(out, r.cg_runs.data(), r.cg_runs.size()); return true; }
  return false;
}

// A number for the metadata JSON: non-finite values as the NaN / Infinity
// tokens Python's json module reads (a degenerate field's isovalue is NaN).
struct JsonNumber {
  double v;
};
inline std::ostream& operator<<(std::ostream& o, JsonNumber n) {
  if (std::isnan(n.v)) return o << "NaN";
  if (std::isinf(n.v)) return o << (n.v > 0 ? "Infinity" : "-Infinity");
  return o << n.v;
}

template <class T> std::string metadata_json(const Result<T>& r) {
  std::ostringstream o;
  o.precision(17);
  o << "{\"precision\": \"" << (sizeof(T) == 4 ? "float32" : (r.prm.float_operators ? "mixed" : "float64"))
    << "\", \"basis_counts\": [";
  for (size_t d = 0; d < r.levels.size(); ++d) o << (d ? ", " : "") << r.levels[d].basis_count;
  o << "], \"aggregates\": " << r.G << ", \"normal_records\": " << r.normal_keys.size() << ", \"timings_seconds\": {";
  for (size_t k = 0; k < r.timings.size(); ++k) o << (k ? ", " : "") << "\"" << r.timings[k].first << "\": " << r.timings[k].second;
  o << "}, \"stage_timing_scope\": \"HIP event elapsed time on the caller's stream, including host count reads\"";
  o << ", \"cube_to_model\": [";
  for (int i = 0; i < 4; ++i) {
    o << (i ? ", " : "") << "[";
    for (int j = 0; j < 4; ++j) o << (j ? ", " : "") << JsonNumber{double(r.cube_to_model[4 * i + j])};
    o << "]";
  }
  const Params& p = r.prm;
  o << "], \"parameters\": {\"depth\": " << p.depth << ", \"full_depth\": " << p.full_depth << ", \"base_depth\": " << p.base_depth
    << ", \"solve_depth\": " << p.solve_depth << ", \"kernel_depth\": " << p.kernel_depth
    << ", \"samples_per_node\": " << p.samples_per_node << ", \"scale\": " << p.scale << ", \"point_weight\": " << p.point_weight
    << ", \"exact_interpolation\": " << (p.exact_interpolation ? "true" : "false") << ", \"iters\": " << p.iters
    << ", \"base_v_cycles\": " << p.base_v_cycles << ", \"cg_accuracy\": " << p.cg_accuracy << "}";
  o << ", \"root_cg_cycles\": [";
  for (size_t k = 0; k < r.cg_runs.size(); ++k) o << (k ? ", " : "") << r.cg_runs[k];
  float wait_ms = 0.f;
  if (r.join_begin && hipEventElapsedTime(&wait_ms, r.join_begin, r.join_end) != hipSuccess) wait_ms = 0.f;
  o << "], \"normal_average_wait_seconds\": " << wait_ms * 1e-3 << ", \"relaxation\": [";
  for (size_t k = 0; k < r.relax_stats.size(); ++k) {
    const auto& st = r.relax_stats[k];
    o << (k ? ", " : "") << "{\"depth\": " << st.depth << ", \"rows\": " << st.rows << ", \"nnz\": " << st.nnz
      << ", \"launches\": " << st.launches << ", \"sweeps\": " << p.iters << ", \"seconds\": " << st.seconds
      << ", \"tile\": " << st.tile << ", \"threads_per_colour\": " << st.width << ", \"tiles\": " << st.tiles
      << ", \"tile_entries\": " << st.entries << ", \"tile_max_colour_rows\": " << st.max_colour_rows << "}";
  }
  o << "], \"profile\": \"pinned cascadic degree-1 Neumann\", \"constant_mode\": \""
    << ((p.base_depth == 0 && p.point_weight == 0) ? "reference root mean term" : "reference solve without an added mean term")
    << "\", \"exact_interpolation_behavior\": \""
    << (p.exact_interpolation ? "empty constraints in pinned reference (issue #16)" : "level-clustered approximate constraints")
    << "\", \"iso_value\": " << JsonNumber{double(r.iso)} << "}";
  return o.str();
}

}  // namespace

// ------------------------------------------------------------ one call
void check_call(int64_t n, int dtype, const Params& prm) {
  if (dtype != 0 && dtype != 1)
    throw PsrError(2, "Reconstruction requires matching float32/float64 3D vectors on the same CUDA device");
  if (n <= 0) throw PsrError(2, "points and normals must have equal nonempty 1D shapes");
  if (prm.float_operators && dtype != 1) throw PsrError(2, "mixed precision (float_operators) runs on float64 inputs");
}

PsrResult* reconstruct_one(const void* points, const void* normals, int64_t n, int dtype, const Params& prm, hipStream_t s) {
  auto res = std::make_unique<PsrResult>();
  res->dtype = dtype;
  if (dtype == 0) {
    Result<float>* r = run<float>((const float*)points, (const float*)normals, n, prm, s);
    res->impl = r;
    res->metadata = metadata_json(*r);
    res->level_offsets_dev.from_host(r->level_offsets.data(), r->level_offsets.size(), s);
  } else {
    Result<double>* r = run<double>((const double*)points, (const double*)normals, n, prm, s);
    res->impl = r;
    res->metadata = metadata_json(*r);
    res->level_offsets_dev.from_host(r->level_offsets.data(), r->level_offsets.size(), s);
  }
  return res.release();
}

// Persistent worker streams per device for batched calls (default priority;
// kept so the relaxation graphs cached per stream are reused across calls).
std::vector<hipStream_t> worker_streams(int dev, int count) {
  static std::mutex mu;
  static std::map<int, std::vector<hipStream_t>> streams;
  std::lock_guard<std::mutex> g(mu);
  auto& v = streams[dev];
  while ((int)v.size() < count) {
    hipStream_t s;
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    v.push_back(s);
  }
  return std::vector<hipStream_t>(v.begin(), v.begin() + count);
}

namespace {
template <class T>
__global__ void k_field_evaluate(PsrField f, const T* __restrict__ queries, int64_t m, T* __restrict__ values,
                                 T* __restrict__ grads, int32_t* __restrict__ valid) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const T q[3] = {queries[3 * i], queries[3 * i + 1], queries[3 * i + 2]};
  T v, g[3];
  const int ok = psr_evaluate_field<T>(f, q, &v, grads ? g : nullptr);
  values[i] = v;
  if (grads) { grads[3 * i] = g[0]; grads[3 * i + 1] = g[1]; grads[3 * i + 2] = g[2]; }
  if (valid) valid[i] = ok;
}
// first key of each level d = 0..depth+1 in sorted keys
__global__ void k_level_offsets(const int64_t* __restrict__ keys, int64_t count, int depth, int32_t* __restrict__ out) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d > depth + 1) return;
  const int64_t key = (int64_t)d << 48;
  int64_t lo = 0, hi = count;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < key) lo = mid + 1; else hi = mid;
  }
  out[d] = (int32_t)lo;
}
}  // namespace

extern "C" {

void psr_params_default(PsrParams* p) {
  p->depth = 6;
  p->full_depth = 3;
  p->base_depth = -1;
  p->solve_depth = -1;
  p->kernel_depth = -1;
  p->samples_per_node = 1.5;
  p->scale = 1.1;
  p->point_weight = 2.0;
  p->exact_interpolation = 0;
  p->iters = 8;
  p->base_v_cycles = 1;
  p->cg_accuracy = 1e-3;
  p->keep_intermediates = 1;
  p->float_operators = 0;
  p->relax_tile = 0;
}

int psr_reconstruct(const void* points, const void* normals, int64_t n, int dtype, const PsrParams* params, void* stream,
                    PsrResult** out) {
  try {
    if (!params || !out) throw PsrError(2, "params and out must be non-null");
    Params prm = resolve(*params);
    check_call(n, dtype, prm);
    *out = reconstruct_one(points, normals, n, dtype, prm, (hipStream_t)stream);
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_reconstruct_batch(const void* const* points, const void* const* normals, const int64_t* counts, int batch, int dtype,
                          const PsrParams* params, int params_count, int max_concurrency, void* stream, PsrResult** out) {
  try {
    if (batch < 0) throw PsrError(2, "batch must be nonnegative");
    if (batch == 0) return 0;
    if (!points || !normals || !counts || !params || !out) throw PsrError(2, "batch arrays must be non-null");
    if (params_count != 1 && params_count != batch) throw PsrError(2, "params_count must be 1 or the batch size");
    std::vector<Params> prm(batch);
    for (int i = 0; i < batch; ++i) {
      try {
        prm[i] = resolve(params[params_count == 1 ? 0 : i]);
        check_call(counts[i], dtype, prm[i]);
      } catch (const PsrError& e) {
        throw PsrError(e.code, "cloud " + std::to_string(i) + ": " + e.what());
      }
      out[i] = nullptr;
    }
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    const int workers = std::max(1, std::min(max_concurrency > 0 ? max_concurrency : 8, batch));
    std::vector<hipStream_t> streams = worker_streams(dev, workers);
    // the caller's stream has produced the inputs: every worker waits for it
    hipEvent_t ready;
    HIP_CHECK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(ready, (hipStream_t)stream));
    for (hipStream_t ws : streams) HIP_CHECK(hipStreamWaitEvent(ws, ready, 0));
    std::vector<int> codes(batch, 0);
    std::vector<std::string> messages(batch);
    const auto start = std::chrono::steady_clock::now();
    auto run_cloud = [&](int i, hipStream_t ws) {
      try {
        out[i] = reconstruct_one(points[i], normals[i], counts[i], dtype, prm[i], ws);
        out[i]->sync_on_free = true;
        HIP_CHECK(hipStreamSynchronize(ws));
        // the cloud's completion time since the batch began, seen by the
        // host after its stream drained (the metadata's last key)
        const double done = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        std::ostringstream o;
        o.precision(9);
        o << ", \"batch_completion_seconds\": " << done << "}";
        std::string& m = out[i]->metadata;
        if (!m.empty() && m.back() == '}') m.replace(m.size() - 1, 1, o.str());
      } catch (const PsrError& e) {
        codes[i] = e.code;
        messages[i] = e.what();
      } catch (const std::exception& e) {
        codes[i] = 1;
        messages[i] = e.what();
      }
    };
    // By default the workers take clouds from a shared counter, so one
    // cloud's latency-bound relaxation overlaps other clouds' stages.  With
    // PSR_BATCH_GROUPS=1 (2 to 8 workers) the clouds, deepest and then
    // largest first, form groups of `workers`; worker w runs member w of
    // every group, and a group's members relax each finer level in one
    // launch sequence, in lockstep.
    const char* env = std::getenv("PSR_BATCH_GROUPS");
    const bool grouped = workers > 1 && workers <= 8 && env && std::string(env) == "1";
    std::vector<int> order(batch);
    for (int i = 0; i < batch; ++i) order[i] = i;
    std::vector<std::unique_ptr<RelaxGroup, void (*)(RelaxGroup*)>> groups;
    if (grouped) {
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (prm[a].depth != prm[b].depth) return prm[a].depth > prm[b].depth;
        return counts[a] > counts[b];
      });
      for (int g = 0; g * workers < batch; ++g)
        groups.emplace_back(relax_group_create(std::min(workers, batch - g * workers)), relax_group_destroy);
    }
    std::atomic<int> next(0);
    auto work = [&](int w) {
      const bool device_ok = hipSetDevice(dev) == hipSuccess;
      if (grouped) {
        for (size_t g = 0; g < groups.size(); ++g) {
          const int idx = (int)g * workers + w;
          if (idx >= batch) break;
          relax_group_bind(groups[g].get(), w);
          if (device_ok) {
            run_cloud(order[idx], streams[w]);
          } else {
            codes[order[idx]] = 1;
            messages[order[idx]] = "hipSetDevice failed on a worker thread";
          }
          relax_group_release();  // also when the cloud failed: the group stops waiting for it
        }
        return;
      }
      if (!device_ok) return;
      for (int i = next++; i < batch; i = next++) run_cloud(i, streams[w]);
    };
    std::vector<std::thread> threads;
    for (int w = 1; w < workers; ++w) threads.emplace_back(work, w);
    work(0);
    for (auto& t : threads) t.join();
    (void)hipEventDestroy(ready);
    for (int i = 0; i < batch; ++i)
      if (codes[i]) {
        for (int j = 0; j < batch; ++j) {
          psr_result_free(out[j]);
          out[j] = nullptr;
        }
        throw PsrError(codes[i], "cloud " + std::to_string(i) + ": " + messages[i]);
      }
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_evaluate(const PsrResult* result, const void* queries, int64_t m, void* values, void* grads, int32_t* valid,
                 void* stream) {
  try {
    if (!result) throw PsrError(2, "null result");
    hipStream_t s = (hipStream_t)stream;
    if (result->dtype == 0) {
      auto* r = (Result<float>*)result->impl;
      float L[9], t[3];
      for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) L[3 * i + j] = r->model_to_cube[4 * i + j];
        t[i] = r->model_to_cube[4 * i + 3];
      }
      evaluate_field(*r, (const float*)queries, m, L, t, (float*)values, (float*)grads, valid, s);
    } else {
      auto* r = (Result<double>*)result->impl;
      double L[9], t[3];
      for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) L[3 * i + j] = r->model_to_cube[4 * i + j];
        t[i] = r->model_to_cube[4 * i + 3];
      }
      evaluate_field(*r, (const double*)queries, m, L, t, (double*)values, (double*)grads, valid, s);
    }
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_result_iso(const PsrResult* result, double* iso) {
  if (!result || !iso) return 2;
  *iso = result->dtype == 0 ? double(((Result<float>*)result->impl)->iso) : ((Result<double>*)result->impl)->iso;
  return 0;
}

int psr_result_field(const PsrResult* result, PsrField* f) {
  if (!result || !f) return 2;
  auto fill = [&](auto* r) {
    f->keys = r->keys.data();
    f->coefficients = r->coefficients.data();
    f->cumulative = r->cumulative.data();
    f->level_offsets = result->level_offsets_dev.data();
    f->count = (int64_t)r->keys.size();
    f->depth = r->prm.depth;
    f->dtype = result->dtype;
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) f->linear[3 * i + j] = double(r->model_to_cube[4 * i + j]);
      f->translation[i] = double(r->model_to_cube[4 * i + 3]);
    }
    f->iso_value = double(r->iso);
  };
  if (result->dtype == 0) fill((Result<float>*)result->impl);
  else fill((Result<double>*)result->impl);
  return 0;
}

int psr_result_metadata_json(const PsrResult* result, char* buf, size_t size, size_t* needed) {
  if (!result) return 2;
  const std::string& m = result->metadata;
  if (needed) *needed = m.size() + 1;
  if (buf && size) {
    size_t k = std::min(size - 1, m.size());
    std::memcpy(buf, m.data(), k);
    buf[k] = 0;
  }
  return 0;
}

int psr_result_array(PsrResult* result, const char* name, int level, const void** data, int64_t* nbytes) {
  try {
    if (!result || !name || !data || !nbytes) throw PsrError(2, "null argument");
    std::string key = std::string(name) + "#" + std::to_string(level);
    auto it = result->cache.find(key);
    if (it == result->cache.end()) {
      std::vector<char> buf;
      bool ok = result->dtype == 0 ? export_array(*(Result<float>*)result->impl, name, level, buf)
                                   : export_array(*(Result<double>*)result->impl, name, level, buf);
      if (!ok) throw PsrError(2, std::string("unknown stage array ") + name);
      it = result->cache.emplace(key, std::move(buf)).first;
    }
    *data = it->second.data();
    *nbytes = (int64_t)it->second.size();
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

void psr_result_free(PsrResult* result) {
  if (!result) return;
  if (result->sync_on_free) (void)hipDeviceSynchronize();
  if (result->dtype == 0) delete (Result<float>*)result->impl;
  else delete (Result<double>*)result->impl;
  delete result;
}

int psr_field_evaluate(const PsrField* field, const void* queries, int64_t m, void* values, void* grads, int32_t* valid,
                       void* stream) {
  try {
    if (!field || !field->keys || !field->level_offsets) throw PsrError(2, "field must be a psr_result_field view");
    if (field->dtype != 0 && field->dtype != 1) throw PsrError(3, "field dtype must be 0 (float32) or 1 (float64)");
    if (m < 0) throw PsrError(2, "query count must be nonnegative");
    if (m == 0) return 0;
    if (!queries || !values) throw PsrError(2, "queries and values must be non-null");
    hipStream_t s = (hipStream_t)stream;
    if (field->dtype == 0)
      hipLaunchKernelGGL(k_field_evaluate<float>, dim3(blocks_for(m, 256)), dim3(256), 0, s, *field, (const float*)queries, m,
                         (float*)values, (float*)grads, valid);
    else
      hipLaunchKernelGGL(k_field_evaluate<double>, dim3(blocks_for(m, 256)), dim3(256), 0, s, *field,
                         (const double*)queries, m, (double*)values, (double*)grads, valid);
    KERNEL_CHECK();
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_field_vjp_inputs(PsrResult* result, const void* queries, int64_t m, const void* value_seed, const void* grad_seed,
                         double* coef_grad, double* cum_grad, void* query_grad, double* transform_grad, void* stream) {
  try {
    if (!result) throw PsrError(2, "null result");
    if (m < 0) throw PsrError(2, "query count must be nonnegative");
    if (m == 0 || (!value_seed && !grad_seed)) return 0;
    if (!queries) throw PsrError(2, "queries must be non-null");
    hipStream_t s = (hipStream_t)stream;
    if (result->dtype == 0)
      field_vjp(*(Result<float>*)result->impl, (const float*)queries, m, (const float*)value_seed, (const float*)grad_seed,
                coef_grad, cum_grad, (float*)query_grad, transform_grad, s);
    else
      field_vjp(*(Result<double>*)result->impl, (const double*)queries, m, (const double*)value_seed, (const double*)grad_seed,
                coef_grad, cum_grad, (double*)query_grad, transform_grad, s);
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_field_vjp(PsrResult* result, const void* queries, int64_t m, const void* value_seed, const void* grad_seed,
                  double* coef_grad, double* cum_grad, void* query_grad, void* stream) {
  return psr_field_vjp_inputs(result, queries, m, value_seed, grad_seed, coef_grad, cum_grad, query_grad, nullptr, stream);
}

int psr_backward(PsrResult* result, const double* coef_grad, const double* cum_grad, double iso_seed, const void* normals,
                 int64_t n, void* grad_normals, void* stream) {
  try {
    if (!result) throw PsrError(2, "null result");
    if (!normals || !grad_normals) throw PsrError(2, "normals and grad_normals must be non-null");
    hipStream_t s = (hipStream_t)stream;
    if (result->dtype == 0) {
      auto* r = (Result<float>*)result->impl;
      if (n != r->n) throw PsrError(2, "normals must be the reconstruction's input normals");
      backward_normals(*r, coef_grad, cum_grad, iso_seed, (float*)grad_normals, (const float*)normals, n, s);
    } else {
      auto* r = (Result<double>*)result->impl;
      if (n != r->n) throw PsrError(2, "normals must be the reconstruction's input normals");
      backward_normals(*r, coef_grad, cum_grad, iso_seed, (double*)grad_normals, (const double*)normals, n, s);
    }
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_backward_inputs(PsrResult* result, const double* coef_grad, const double* cum_grad, double iso_seed,
                        const double* transform_grad, const void* points, const void* normals, int64_t n, void* grad_points,
                        void* grad_normals, void* stream) {
  try {
    if (!result) throw PsrError(2, "null result");
    if (!grad_points && !grad_normals) throw PsrError(2, "grad_points and grad_normals are both null");
    if (grad_points && !points) throw PsrError(2, "grad_points needs the input points");
    if (grad_normals && !normals) throw PsrError(2, "grad_normals needs the input normals");
    hipStream_t s = (hipStream_t)stream;
    auto run = [&](auto* r, auto tag) {
      using T = decltype(tag);
      if (n != r->n) throw PsrError(2, "points and normals must be the reconstruction's inputs");
      backward_inputs(*r, coef_grad, cum_grad, iso_seed, transform_grad, (const T*)points, (const T*)normals, n, (T*)grad_points,
                      (T*)grad_normals, s);
    };
    if (result->dtype == 0)
      run((Result<float>*)result->impl, 0.0f);
    else
      run((Result<double>*)result->impl, 0.0);
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_relax_sliced(int dtype, const int64_t* keys, int64_t rows, const int32_t* offsets, const int32_t* columns,
                     const void* values, const void* b, void* x, int depth, int iters, void* stream) {
  try {
    if (depth < 0 || depth > 15 || iters < 1) throw PsrError(2, "Require 0 <= depth <= 15 and iters >= 1");
    if (rows < 0 || rows > INT32_MAX) throw PsrError(2, "row count out of range");
    hipStream_t s = (hipStream_t)stream;
    if (dtype == 0)
      relax_sliced_csr<float>(keys, (int)rows, offsets, columns, (const float*)values, (const float*)b, (float*)x, depth, iters, s);
    else if (dtype == 1)
      relax_sliced_csr<double>(keys, (int)rows, offsets, columns, (const double*)values, (const double*)b, (double*)x, depth,
                               iters, s);
    else
      throw PsrError(3, "dtype must be 0 (float32) or 1 (float64)");
    HIP_CHECK(hipStreamSynchronize(s));
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_relax_scheduled(int dtype, const int64_t* keys, int64_t rows, const int32_t* offsets, const int32_t* columns,
                        const void* values, const void* b, void* x, int depth, int iters, int forward, int z, void* stream) {
  try {
    if (depth < 0 || depth > 15 || iters < 1) throw PsrError(2, "Require 0 <= depth <= 15 and iters >= 1");
    if (rows < 0 || rows > INT32_MAX) throw PsrError(2, "row count out of range");
    hipStream_t s = (hipStream_t)stream;
    if (dtype == 0)
      relax_scheduled_csr<float>(keys, (int)rows, offsets, columns, (const float*)values, (const float*)b, (float*)x, depth, iters,
                                 forward, z, s);
    else if (dtype == 1)
      relax_scheduled_csr<double>(keys, (int)rows, offsets, columns, (const double*)values, (const double*)b, (double*)x, depth,
                                  iters, forward, z, s);
    else
      throw PsrError(3, "dtype must be 0 (float32) or 1 (float64)");
    HIP_CHECK(hipStreamSynchronize(s));
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

// ---------------------------------------------------------------- marching cubes (meshing.py)
struct PsrMesh {
  McState* state = nullptr;
};

int psr_mesh_begin(const int32_t shape[3], const float lower[3], const float upper[3], void* stream, PsrMesh** out) {
  try {
    if (!shape || !lower || !upper || !out) throw PsrError(2, "null argument");
    int sh[3] = {shape[0], shape[1], shape[2]};
    auto* m = new PsrMesh();
    try {
      m->state = mc_create(sh, lower, upper, (hipStream_t)stream);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_mesh_slab_nodes(PsrMesh* m, int32_t start, int32_t width, int64_t* nodes) {
  try {
    if (!m || !nodes) throw PsrError(2, "null argument");
    *nodes = mc_slab_nodes(*m->state, start, width);
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_mesh_slab_points(PsrMesh* m, int32_t start, int32_t width, float* points) {
  try {
    if (!m || !points) throw PsrError(2, "null argument");
    mc_slab_points(*m->state, start, width, points);
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_mesh_slab_values(PsrMesh* m, int32_t start, int32_t width, const float* values) {
  try {
    if (!m || !values) throw PsrError(2, "null argument");
    mc_slab_values(*m->state, start, width, values);
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_mesh_select(PsrMesh* m, int64_t* retained, int32_t* invalid) {
  try {
    if (!m || !retained || !invalid) throw PsrError(2, "null argument");
    int bad = 0;
    *retained = mc_select(*m->state, &bad);
    *invalid = bad;
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_mesh_corner_points(PsrMesh* m, float* points) {
  try {
    if (!m) throw PsrError(2, "null argument");
    mc_corner_points(*m->state, points);
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_mesh_triangulate(PsrMesh* m, const float* corner_values, int64_t* vertices, int64_t* triangles) {
  try {
    if (!m || !vertices || !triangles) throw PsrError(2, "null argument");
    int V = 0, T = 0;
    mc_triangulate(*m->state, corner_values, &V, &T);
    *vertices = V;
    *triangles = T;
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_mesh_copy(PsrMesh* m, float* vertices, int32_t* indices) {
  try {
    if (!m) throw PsrError(2, "null argument");
    mc_copy(*m->state, vertices, indices);
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_mesh_vertex_vjp(PsrMesh* m, const float* corner_values, const float* grad_vertices, float* grad_values) {
  try {
    if (!m) throw PsrError(2, "null argument");
    mc_vertex_vjp(*m->state, corner_values, grad_vertices, grad_values);
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_mesh_case_table(int8_t* counts, int8_t* edges) {
  try {
    if (!counts || !edges) throw PsrError(2, "null argument");
    const McTable& t = mc_table();
    std::memcpy(counts, t.count, sizeof(t.count));
    std::memcpy(edges, t.edges, sizeof(t.edges));
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

void psr_mesh_free(PsrMesh* m) {
  if (!m) return;
  mc_free(m->state);
  delete m;
}

// ---------------------------------------------------------------- tree metadata (tree_metadata.py)
struct PsrTreeMetadata {
  TreeMetadata meta;
};

int psr_tree_metadata(PsrResult* result, void* stream, PsrTreeMetadata** out) {
  try {
    if (!result || !out) throw PsrError(2, "null argument");
    auto* m = new PsrTreeMetadata();
    try {
      hipStream_t s = (hipStream_t)stream;
      if (result->dtype == 0) build_tree_metadata(*(Result<float>*)result->impl, m->meta, s);
      else build_tree_metadata(*(Result<double>*)result->impl, m->meta, s);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_tree_metadata_levels(const PsrTreeMetadata* m, int32_t* count) {
  if (!m || !count) return fail(PsrError(2, "null argument"));
  *count = (int32_t)m->meta.levels.size();
  return 0;
}

int psr_tree_metadata_level(const PsrTreeMetadata* m, int32_t index, int32_t* local_depth, int64_t* count) {
  if (!m || !local_depth || !count) return fail(PsrError(2, "null argument"));
  if (index < 0 || index >= (int32_t)m->meta.levels.size()) return fail(PsrError(2, "level index out of range"));
  *local_depth = m->meta.levels[index].local_depth;
  *count = (int64_t)m->meta.levels[index].keys.size();
  return 0;
}

int psr_tree_metadata_copy(const PsrTreeMetadata* m, int32_t index, int64_t* keys, uint8_t* flags, uint8_t* ghost,
                           int32_t* parents, int32_t* children, void* stream) {
  try {
    if (!m) throw PsrError(2, "null argument");
    if (index < 0 || index >= (int32_t)m->meta.levels.size()) throw PsrError(2, "level index out of range");
    const TreeMetadata::Level& lv = m->meta.levels[index];
    const size_t n = lv.keys.size();
    hipStream_t s = (hipStream_t)stream;
    if (n) {
      if (keys) HIP_CHECK(hipMemcpyAsync(keys, lv.keys.data(), sizeof(int64_t) * n, hipMemcpyDeviceToDevice, s));
      if (flags) HIP_CHECK(hipMemcpyAsync(flags, lv.flags.data(), n, hipMemcpyDeviceToDevice, s));
      if (ghost) HIP_CHECK(hipMemcpyAsync(ghost, lv.ghost.data(), n, hipMemcpyDeviceToDevice, s));
      if (parents) HIP_CHECK(hipMemcpyAsync(parents, lv.parents.data(), sizeof(int) * n, hipMemcpyDeviceToDevice, s));
      if (children) HIP_CHECK(hipMemcpyAsync(children, lv.children.data(), sizeof(int) * 8 * n, hipMemcpyDeviceToDevice, s));
    }
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

void psr_tree_metadata_free(PsrTreeMetadata* m) { delete m; }

int psr_field_level_offsets(const int64_t* keys, int64_t count, int depth, int32_t* offsets, void* stream) {
  try {
    if (depth < 0 || depth > 15) throw PsrError(2, "Keys support depths 0..15, including upper boundary nodes");
    if (count < 0 || count > INT32_MAX) throw PsrError(2, "key count must lie in [0, 2**31)");
    if (!offsets || (count && !keys)) throw PsrError(2, "null argument");
    hipLaunchKernelGGL(k_level_offsets, dim3(1), dim3(64), 0, (hipStream_t)stream, keys, count, depth, offsets);
    KERNEL_CHECK();
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_isovalue(int dtype, const void* values, const void* weights, int64_t n, void* out, void* stream) {
  try {
    if (dtype != 0 && dtype != 1) throw PsrError(3, "dtype must be 0 (float32) or 1 (float64)");
    if (n < 0) throw PsrError(2, "value count must be nonnegative");
    if (!out || (n && (!values || !weights))) throw PsrError(2, "null argument");
    hipStream_t s = (hipStream_t)stream;
    if (dtype == 0) compute_isovalue((const float*)values, (const float*)weights, n, (float*)out, s);
    else compute_isovalue((const double*)values, (const double*)weights, n, (double*)out, s);
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

int psr_isovalue_noalloc(int dtype, const void* values, const void* weights, int64_t n, void* out, void* stream) {
  try {
    if (dtype != 0 && dtype != 1) throw PsrError(3, "dtype must be 0 (float32) or 1 (float64)");
    if (n < 0) throw PsrError(2, "value count must be nonnegative");
    if (!out || (n && (!values || !weights))) throw PsrError(2, "null argument");
    hipStream_t s = (hipStream_t)stream;
    if (dtype == 0) isovalue_noalloc((const float*)values, (const float*)weights, n, (float*)out, s);
    else isovalue_noalloc((const double*)values, (const double*)weights, n, (double*)out, s);
    return 0;
  } catch (const PsrError& e) {
    return fail(e);
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

const char* psr_last_error(void) { return g_error.c_str(); }

void psr_release_cache(void) { Pool::instance().release_all(); }

}  // extern "C"
