// Device-resident reconstruction state.  One Result owns every device array
// of one reconstruction (keys, coefficients, cumulative coefficients and the
// per-stage intermediates kept for diagnostics and parity checks).
#pragma once
#include "psr_runtime.h"

namespace psr {

struct Params {
  int depth = 6, full_depth = 3, base_depth = 3, solve_depth = 6, kernel_depth = 4;
  double samples_per_node = 1.5, scale = 1.1, point_weight = 2.0;
  int exact_interpolation = 0, iters = 8, base_v_cycles = 1;
  double cg_accuracy = 1e-3;
  int keep_intermediates = 1;  // keep per-stage arrays for diagnostics
  int float_operators = 0;     // finer-level operators stored as float (double runs)
  int relax_tile = 0;          // tiled relaxation side (0 default, -1 colour-per-launch form)
};

// 27-slot stencil matrix (slot k <-> neighbour offset (k/9-1, k/3%3-1, k%3-1)
// around the row node or its parent), slot-major: cols[k*rows + i].  Slot
// order equals the ascending-column order of the reference's sorted CSR rows.
template <class V> struct Stencil {
  int rows = 0, cols_count = 0;
  DArray<int> cols;
  DArray<V> vals;
};

template <class T> struct LevelState {
  DArray<int64_t> tree_keys, basis_keys;
  int basis_count = 0;
  ColumnIndex columns;  // (x, y) runs of basis_keys
  DArray<int> dense;    // z-major dense index of basis_keys during the solve stage
  int dense_dim = 0;
  // restriction gathers to this level from level d+1: 27 fine indices per row
  DArray<int> restriction;
  // aggregates (density.py:228-240): keys and (sum x, sum y, sum z, count)
  DArray<int64_t> agg_keys;
  DArray<T> agg_vals;  // 4 per record
  // density splat level (density.py:245-271)
  DArray<int64_t> dens_keys;
  DArray<T> dens_vals;
  // operators and solve
  Stencil<T> A, cross;      // cast to run precision
  Stencil<float> A32, cross32;  // finer levels of mixed-precision runs (instead of A, cross)
  Stencil<T> P;             // prolongation rows (8 slots, x-major corner order)
  DArray<double> normal_load, point_load;
  DArray<T> rhs, correction;
  // finer levels keep A, rhs and point_load in sliced schedule order
  // (position j holds row order[j]); cross keeps canonical rows
  DArray<int> order;
  bool scheduled = false;
  bool solved = false;
};

template <class T> struct Result {
  Params prm;
  hipStream_t s = nullptr;
  int64_t n = 0;
  int G = 0;
  // samples (samples.py)
  T bounds[6];
  T transform[5];
  DArray<T> q, unit_normals;
  DArray<int> input_status, input_group, input_order;
  DArray<int64_t> sample_keys;
  DArray<T> position_sums, normal_sums, weights;
  DArray<int> first_indices, group_offsets;
  // density (density.py)
  T density_scale;
  DArray<T> densities, dw;
  // normals (normals.py)
  DArray<int64_t> normal_keys;
  DArray<T> normal_vals;  // 3 per record (already negated)
  std::vector<int> normal_offsets;
  DArray<T> normal_average;  // 3 values on device
  T normal_average_host[3];
  DArray<int> normal_order;
  // The serial normal average runs on a side stream, overlapped with the
  // tree stage; its input stays alive until the main stream joins it.
  hipStream_t side = nullptr;  // a companion stream kept for the process (not owned)
  hipEvent_t average_done = nullptr;
  bool average_pending = false;
  DArray<T> average_contrib;
  // HIP-event span of the main stream's wait for the side stream (metadata)
  hipEvent_t join_begin = nullptr, join_end = nullptr;
  void join_average() {
    if (!average_pending) return;
    if (!join_begin) {
      HIP_CHECK(hipEventCreate(&join_begin));
      HIP_CHECK(hipEventCreate(&join_end));
    }
    HIP_CHECK(hipEventRecord(join_begin, s));
    HIP_CHECK(hipStreamWaitEvent(s, average_done, 0));
    HIP_CHECK(hipEventRecord(join_end, s));
    average_pending = false;
  }
  // tree + solve per level
  std::vector<LevelState<T>> levels;
  std::vector<int> cg_runs;
  DArray<int> cg_ran;  // per base V-cycle: 1 when the root CG ran; read into cg_runs once the run has synchronised
  // root solve inputs per base V-cycle (right-hand side, initial x), kept for
  // the reverse pass
  std::vector<DArray<T>> root_b, root_x;
  // field
  DArray<int64_t> keys;
  DArray<T> coefficients, cumulative, centroid_values;
  std::vector<int> level_offsets;  // concatenated key offsets per depth
  DArray<int> offsets_dev;         // the same on the device
  // (x, y) column runs of every level of keys (absolute indices), laid out
  // level after level with (2^d + 1)^2 cells each; empty above depth 11
  DArray<int> column_start, column_end;
  T iso;
  T cube_to_model[16], model_to_cube[16];
  // timings (seconds) per stage, from HIP events
  std::vector<std::pair<std::string, float>> timings;
  // sliced relaxation per finer level: rows, stored nonzeros, kernel launches
  // and the HIP-event span of its launch sequence on the stream
  struct RelaxStat {
    int depth, rows;
    int64_t nnz;
    int launches;
    hipEvent_t start, end;
    float seconds;
    // tiled form: tile side, threads per colour, tiles and tile entries
    // (0 for the colour-per-launch form)
    int tile = 0, width = 0, tiles = 0, entries = 0, max_colour_rows = 0;
  };
  std::vector<RelaxStat> relax_stats;
  // device nonzero counters of the relaxed levels, read once the run is done
  std::vector<DArray<unsigned long long>> relax_nnz;
  ~Result() {
    if (join_begin) {
      (void)hipEventSynchronize(join_end);
      (void)hipEventDestroy(join_begin);
      (void)hipEventDestroy(join_end);
    }
    if (average_done) {
      (void)hipEventSynchronize(average_done);
      (void)hipEventDestroy(average_done);
    }
    for (auto& st : relax_stats) {
      (void)hipEventDestroy(st.start);
      (void)hipEventDestroy(st.end);
    }
  }
};

// stage entry points (one translation unit each)
template <class T> void stage_samples(Result<T>& r, const T* pts, const T* nrm);
template <class T> void stage_density(Result<T>& r);
template <class T> void stage_normals(Result<T>& r);
template <class T> void stage_tree(Result<T>& r);
// Field reverse mode (psr_adjoint.hip): value / model-gradient seeds at
// model-space queries -> float64 coefficient and cumulative gradients
// (atomically accumulated) and query gradients (accumulated, may be null).
template <class T>
void field_vjp(Result<T>& r, const T* queries, int64_t m, const T* gvalue, const T* ggrad, double* gcoef, double* gcum,
               T* gquery, double* gtransform, hipStream_t s);
// Coefficient / cumulative / isovalue seeds -> input-normal gradients
// (accumulated into gnormals, n x 3).
template <class T>
void backward_normals(Result<T>& r, const double* gcoef, const double* gcum, double iso_seed, T* gnormals, const T* normals,
                      int64_t n, hipStream_t s);
// The same seeds (plus gtransform[4], the loss's gradient w.r.t. the
// normalisation (inv, t1, t2, t3) through the query mapping, or null) ->
// gradients w.r.t. the input positions (gpoints, runs without screening)
// and/or normals (gnormals); either may be null.  Both are accumulated (n x 3).
template <class T>
void backward_inputs(Result<T>& r, const double* gcoef, const double* gcum, double iso_seed, const double* gtransform,
                     const T* points, const T* normals, int64_t n, T* gpoints, T* gnormals, hipStream_t s);
// Matrix-side gradients of a solve, per level d in [base, top]: w.r.t. the
// operator values (gA: the level's storage layout, schedule positions for
// finer levels and canonical rows for the base level), the finer levels'
// cross operators (gC, canonical rows) and the right-hand sides (grhs,
// canonical).
struct SolveGrads {
  int top = -1;
  std::vector<DArray<double>> gA, gC, grhs;
};
// Reverse of the cascadic solve and normal loads: gcoef (float64, one per
// concatenated coefficient) gives float64 gradients of the normal records'
// three values (accumulated into grecords); with mg, also the matrix-side
// gradients (the relaxations are replayed once for the x they saw).
template <class T>
void solve_backward(Result<T>& r, const double* gcoef, double* grecords, hipStream_t s, SolveGrads* mg = nullptr);
// Screening reversed (assembly_adjoint.py:44-91, constraints.py:55-84): the
// operator, cross-operator and point-load gradients of mg -> the constraint
// points' aggregates (gagg[d], 3 per aggregate of level d, accumulated) and
// the normal average's weight component (gavg1, device scalar, accumulated).
template <class T>
void screening_backward(Result<T>& r, const SolveGrads& mg, std::vector<DArray<double>>& gagg, double* gavg1, hipStream_t s);
template <class T> void stage_solve(Result<T>& r);
// reconstruction.py:101-116 (_compute_isovalue): *out (device) = the
// weighted mean sum(double(v*w)) / sum(double(w)) of G aggregates, as one
// serial double chain for G <= 1024, else double partials of 1024-element
// tiles (256 lanes, elements lane + 256k in order, then a halving tree)
// combined the same way.  isovalue_launch takes the partial buffers (tp, mp:
// isovalue_partials(G) doubles each); compute_isovalue allocates them.
inline int64_t isovalue_partials(int64_t G) { return (G + 1023) / 1024; }
template <class T> void isovalue_launch(const T* v, const T* w, int64_t G, double* tp, double* mp, T* out, hipStream_t s);
template <class T> void compute_isovalue(const T* v, const T* w, int64_t G, T* out, hipStream_t s);
// The same reduction with its partials in a file-scope device array of
// 2 * ISOVALUE_NOALLOC_PARTIALS doubles (1 MiB per GPU): no allocation; at
// most ISOVALUE_NOALLOC_MAX values; calls that overlap on one GPU race.
constexpr int64_t ISOVALUE_NOALLOC_MAX = (int64_t)1 << 26;
constexpr int64_t ISOVALUE_NOALLOC_PARTIALS = ISOVALUE_NOALLOC_MAX / 1024;
template <class T> void isovalue_noalloc(const T* v, const T* w, int64_t G, T* out, hipStream_t s);
template <class T> void stage_field(Result<T>& r);
template <class T>
void evaluate_field(const Result<T>& r, const T* queries, int64_t m, const T* linear, const T* translation, T* values,
                    T* grads, int* valid, hipStream_t s);

template <class T>
void relax_sliced_csr(const int64_t* keys, int R, const int* offsets, const int* columns, const T* values, const T* b, T* x,
                      int depth, int iters, hipStream_t s);

template <class T>
void relax_scheduled_csr(const int64_t* keys, int R, const int* offsets, const int* columns, const T* values, const T* b, T* x,
                         int depth, int iters, int forward, int z, hipStream_t s);

// shared device helpers used by several stages
template <class T> void sort_records_u64(DArray<uint64_t>& keys, DArray<int>& ids, int64_t n, int bits, hipStream_t s);

// Host-side transform (samples.py:54-88) -- scalar work done once per call.
// Batched reconstructions relax in groups of up to 8: a worker thread binds
// itself to a member slot for one reconstruction and releases it after (also
// on failure); see RelaxGroup in psr_solve.hip.
struct RelaxGroup;
RelaxGroup* relax_group_create(int members);
void relax_group_destroy(RelaxGroup* g);
void relax_group_bind(RelaxGroup* g, int member);
void relax_group_release();

template <class T> int compute_transform(const T* b, double scale, T* t);
template <class T> void reference_inverse(const T* m, T* out);

}  // namespace psr
