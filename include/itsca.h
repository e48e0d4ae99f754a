/* itsca.h -- C ABI of libitsca_hip.so, the gfx950 cellular-automaton track
 * finder of the ALICE ITS (plain CA: index tables, tracklets, cells,
 * neighbours, CA levels and roads).
 *
 * The reference plugs its back ends in at compile time, through
 * TrackerTraits<IsGPU>::computeLayerTracklets / computeLayerCells
 * (include/ITSReconstruction/CA/Tracker.h:39-48) with the device state hung off
 * PrimaryVertexContext (PrimaryVertexContext.h:59-77,
 * gpu/PrimaryVertexContext.h:39-42).  This library replaces that whole back
 * end -- and the stages the reference keeps on the host (findCellsNeighbours,
 * findTracks, computeMontecarloLabels, Tracker.cxx:416-652) -- behind plain C
 * entry points.  The C++ mirror of the reference API (Tracker<true>) and any
 * foreign-function binding sit on top of it.
 *
 * Conventions: every call returns ITSCA_OK (0) or an error code, with the text
 * in itsca_last_error(ctx).  Device memory is owned by the context; host
 * buffers passed in are copied before the call returns; output buffers are
 * owned by the caller, sized from itsca_get_counts().  A context is not
 * thread-safe; independent contexts are.
 */
#ifndef ITSCA_H
#define ITSCA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct itsca_ctx itsca_ctx;
typedef int itsca_status;

enum {
  ITSCA_OK = 0,
  ITSCA_ERR_ARGUMENT = 1,
  ITSCA_ERR_HIP = 2,      /* a HIP runtime call failed (no device, launch failure, ...) */
  ITSCA_ERR_CAPACITY = 3, /* an output did not fit even after growing the buffers */
  ITSCA_ERR_STATE = 4     /* call out of order (e.g. download before run) */
};

typedef struct itsca_config {
  int device;            /* HIP device ordinal */
  int persistent_blocks; /* workgroups of the persistent stage kernels (some take at most their resident
                            workgroups, two take twice as many); 0 = 8 per CU */
  double capacity_scale; /* multiplier on the built-in output-size estimates; 0 = 1 */
} itsca_config;

/* Output sizes of one work item (one primary vertex of one event). */
typedef struct itsca_counts {
  int clusters[7];
  int tracklets[6];
  int cells[5];
  int neighbours[4]; /* accepted edges between cell layers l and l+1 */
  int roads;
} itsca_counts;

/* Record kinds for itsca_download (rows of 4-byte fields, little endian). */
enum {
  ITSCA_CLUSTERS = 0,          /* sorted clusters of `layer`: id, phi(f32), r(f32), bin, x, y, z (f32), mc */
  ITSCA_INDEX_TABLE = 1,       /* 401 int: first sorted position of each z-phi bin of `layer` (1..6) */
  ITSCA_TRACKLETS = 2,         /* first, second, tanLambda(f32), phi(f32) */
  ITSCA_TRACKLET_LOOKUP = 3,   /* per cluster of layer+1: first tracklet of layer+1 starting there, or -1 */
  ITSCA_CELLS = 4,             /* c1, c2, c3, t1, t2, nx, ny, nz, curvature (f32), level */
  ITSCA_CELL_LOOKUP = 5,       /* per tracklet of layer+1: first cell of layer+1 starting there, or -1 */
  ITSCA_NEIGHBOUR_OFFSETS = 6, /* cells(layer+1)+1 int: start of each cell's list in ITSCA_NEIGHBOURS */
  ITSCA_NEIGHBOURS = 7,        /* ascending cells of `layer` neighbouring each cell of layer+1 */
  ITSCA_ROADS = 8              /* cell index per layer 0..4, size, label, fake (8 int) */
};

/* Creates a context on cfg->device (cfg may be NULL: device 0, defaults).  On
 * failure *out still receives a context that holds the error text for
 * itsca_last_error; destroy it as usual. */
itsca_status itsca_create(itsca_ctx** out, const itsca_config* cfg);
void itsca_destroy(itsca_ctx* ctx);
const char* itsca_last_error(const itsca_ctx* ctx);

/* Input: events are staged on the host, then uploaded to HBM together.
 * Replaces Event/Layer::addCluster + PrimaryVertexContext::initialize's input
 * (Event.h:35-44, PrimaryVertexContext.cxx:31-50).  Arrays hold every layer's
 * clusters, layer after layer, in file order; layer_offsets has 8 entries
 * (layer l occupies [layer_offsets[l], layer_offsets[l+1])); vertices holds 3
 * floats per primary vertex. */
itsca_status itsca_clear_events(itsca_ctx* ctx);
itsca_status itsca_add_event(itsca_ctx* ctx, const float* x, const float* y, const float* z,
                             const int32_t* cluster_id, const int32_t* mc_id, const int32_t* layer_offsets,
                             const float* vertices, int num_vertices, int* event_index);
itsca_status itsca_upload(itsca_ctx* ctx);

/* Runs the plain CA on every (event, vertex) work item of the uploaded batch
 * and waits for it.  Counterpart of Tracker::clustersToTracks
 * (Tracker.cxx:244-265) for the whole batch; work item i is the i-th vertex in
 * event order. */
itsca_status itsca_run(itsca_ctx* ctx);

/* Enqueues the same run on the context's stream without waiting; call
 * itsca_finish before reading results.  Used to time back-to-back runs. */
itsca_status itsca_run_async(itsca_ctx* ctx);
itsca_status itsca_finish(itsca_ctx* ctx);

int itsca_num_items(const itsca_ctx* ctx);

/* The roads of every work item of the last run in one copy: rows of 8 int
 * (the ITSCA_ROADS format), item after item; first_road[i] is the first row
 * of item i and first_road[num_items] the total (first_road holds
 * num_items + 1 entries).  cap_rows is the capacity of rows; *total receives
 * the number of roads (ITSCA_ERR_ARGUMENT, nothing copied, when it exceeds
 * cap_rows).  The batch counterpart of the roads that
 * Tracker::clustersToTracks returns by value (Tracker.cxx:244-265). */
itsca_status itsca_download_roads(itsca_ctx* ctx, int32_t* rows, size_t cap_rows, int32_t* first_road,
                                  size_t* total);

/* Output capacities of the device buffers after the last run (elements, all
 * work items and layers together); they grow on overflow and never shrink.
 * Counterpart of the capacities the reference reserves per layer
 * (PrimaryVertexContext.cxx:56-74), reported by its memory benchmark
 * (Tracker.cxx:303-366). */
typedef struct itsca_capacities {
  long tracklets;
  long cells;
  long neighbours;
  long roads;
} itsca_capacities;
itsca_status itsca_get_capacities(itsca_ctx* ctx, itsca_capacities* out);
itsca_status itsca_get_counts(itsca_ctx* ctx, int item, itsca_counts* out);

/* Copies one record kind of one work item into dst (cap_bytes long);
 * *written receives the number of bytes. */
itsca_status itsca_download(itsca_ctx* ctx, int item, int what, int layer, void* dst, size_t cap_bytes,
                            size_t* written);

/* Launch shape of the stage kernels and binning of the tracklet search index.
 * None of it changes a result, only the time (itsca_autotune searches it).
 *   block_size:        workgroup size and chunk length, 64, 128 or 256
 *   persistent_blocks: workgroups of the persistent stage kernels (0 = 8 per CU); the tracklet count, the
 *     scoring, the cell fill, the edge count and the levels launch at most their resident workgroups, the
 *     edge fill and the road fill twice this number
 *   z_subbins:         z sub-bins per z bin of the search index, 1..8 (0 = 8);
 *                      the search tests 20 * z_subbins z slices per layer instead
 *                      of the reference's 20 bins
 *   phi_subrows:       phi sub-rows per phi row of the search index, 1, 2 or 4
 *                      (0 = 1); the search sees 20 * phi_subrows phi slices per
 *                      layer instead of the reference's 20 rows.
 *                      z_subbins * phi_subrows is at most 15.
 * The reference's own 20 x 20 binning (Constants.h:77-79) is kept for the
 * results: it decides which candidate its inclusive bound adds
 * (Tracker.cxx:84-89), so the sub-bins only index the search. */
typedef struct itsca_tuning {
  int block_size;
  int persistent_blocks;
  int z_subbins;
  int phi_subrows;
} itsca_tuning;
itsca_status itsca_set_tuning(itsca_ctx* ctx, const itsca_tuning* tuning);
itsca_status itsca_get_tuning(itsca_ctx* ctx, itsca_tuning* tuning);

/* Autotuner: times every tuning (block size 64/128/256 x 4/8/16 workgroups
 * per CU x search-index binnings: z sub-bins 1/2/4/8 x phi sub-rows 1/2/4,
 * at most 15 sub-bins per bin) over `repeats` runs of the uploaded batch,
 * keeps the fastest, and leaves the results of one run with it.  *best_ms gets
 * its time per run. */
itsca_status itsca_autotune(itsca_ctx* ctx, int repeats, itsca_tuning* chosen, float* best_ms);

/* Runs the pipeline as one captured HIP graph (captured on first use, replayed
 * afterwards, captured again after the buffers grow or a new batch is
 * uploaded).  Results are identical to the launch-by-launch run; profiled
 * runs are never captured. */
itsca_status itsca_set_graph(itsca_ctx* ctx, int enable);

/* ---------------------------------------------------------------------------
 * One event without device memory allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture elsewhere in the program, tight serving loops, harnesses that
 * forbid hipMalloc).  Runs the same pipeline as itsca_run (with the context's
 * tuning and learned-filter setting) on one event with one primary vertex
 * whose clusters the caller already holds in device memory, and writes the
 * roads to the caller's device buffer.  Every intermediate array lives in a
 * static device array of ITSCA_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; the call allocates nothing and waits for its results.
 *   - up to ITSCA_NOALLOC_MAX_CLUSTERS clusters (else ITSCA_ERR_ARGUMENT); an
 *     intermediate stage that outgrows the scratch gives ITSCA_ERR_CAPACITY
 *     (there is no growing);
 *   - x, y, z, cluster_id, mc_id: device arrays; layer l occupies
 *     [layer_offsets[l], layer_offsets[l+1]) of them; vertex: host values;
 *   - roads: device buffer of roads_cap rows of 8 int in the ITSCA_ROADS
 *     format; *num_roads gets the number of roads, and ITSCA_ERR_CAPACITY is
 *     returned (with the first roads_cap rows written) when it exceeds roads_cap.
 * Every call on one GPU shares that GPU's scratch: calls that overlap on one
 * GPU (two contexts, two threads) race and must be serialised by the caller.
 * --------------------------------------------------------------------------- */
#define ITSCA_NOALLOC_SCRATCH_BYTES (128u << 20)
#define ITSCA_NOALLOC_MAX_CLUSTERS 131072
typedef struct itsca_device_event {
  const float *x, *y, *z;
  const int32_t *cluster_id, *mc_id;
  int32_t layer_offsets[8];
  float vertex[3];
} itsca_device_event;
itsca_status itsca_run_noalloc(itsca_ctx* ctx, const itsca_device_event* ev, int32_t* roads, int roads_cap,
                               int* num_roads);

/* ---------------------------------------------------------------------------
 * Learned tracklet filter (no counterpart in the reference; an optional stage
 * between the tracklet search and cell building).  A 10 -> 32 -> 32 -> 1 ReLU
 * network with BF16 weights and activations and FP32 accumulation scores every
 * tracklet that passed the exact tracklet cuts; tracklets below the threshold
 * are dropped before cell building, and every cut after that is the exact CA
 * one, so the roads found are always a subset of the plain roads.
 * Parameters: 1441 floats, W1[32][10] b1[32] W2[32][32] b2[32] w3[32] b3.
 * --------------------------------------------------------------------------- */
#define ITSCA_MLP_PARAMS 1441
#define ITSCA_MLP_FEATURES 16 /* BF16 per training row: 10 features, 6 zeros */
itsca_status itsca_mlp_init(itsca_ctx* ctx, uint64_t seed); /* He initialisation */
itsca_status itsca_mlp_set_params(itsca_ctx* ctx, const float* params);
itsca_status itsca_mlp_get_params(itsca_ctx* ctx, float* params);
/* Turns the filter on for later runs; tracklets with logit < threshold are dropped.
 * enable: 0 off; 1 on, scored by a kernel of its own after the search of every
 * tracklet layer; 2 on, scored inside the search (one launch per tracklet
 * layer, no candidate round trip through memory; slower on MI355X).  Both give
 * the same tracklets; form 2 falls back to form 1 for batches searched with
 * more than one z sub-bin or with a layer of more than 65536 clusters. */
itsca_status itsca_set_filter(itsca_ctx* ctx, int enable, float threshold);
/* phi gates of the filtered runs: per tracklet layer l (ITSCA_TRACKLET_LAYERS
 * of them), the tracklet search of a run with the filter on keeps only the
 * candidates with |dphi| < gates[l] (0 < gates[l] <= 0.3, the reference's phi
 * cut: no gate) -- the plain candidates that pass the gate, in the plain
 * order -- and searches only the window rows that can hold them.  The network
 * scores what is left, its count features being the gated counts.  gates ==
 * NULL: no gates (the default).  Plain runs never use them. */
#define ITSCA_TRACKLET_LAYERS 6
itsca_status itsca_set_filter_gates(itsca_ctx* ctx, const float* gates);

/* Training set on the device: rows of ITSCA_MLP_FEATURES BF16 values and a
 * 0/1 label (the tracklet belongs to a cell of a final road of the plain CA). */
itsca_status itsca_train_clear(itsca_ctx* ctx, int capacity);
/* After a plain itsca_run, or a filtered one whose threshold keeps every
 * candidate (<= -1e38: the gated candidates): labels every tracklet of the
 * batch (a tracklet of a cell of a final road of that run) and appends every
 * stride-th one (starting at `phase`) to the training set. */
itsca_status itsca_train_collect(itsca_ctx* ctx, int stride, int phase, int* added);
itsca_status itsca_train_set(itsca_ctx* ctx, const uint16_t* features, const uint8_t* labels, int n);
int itsca_train_size(const itsca_ctx* ctx);
itsca_status itsca_train_get(itsca_ctx* ctx, uint16_t* features, uint8_t* labels, int n);
/* One step on rows (mult*k + offset) mod n, k in [first, first+count): writes the
 * gradient of the weighted binary cross-entropy (mean over the sample weights;
 * positives weigh pos_weight) and, when lr > 0, applies AdamW (beta1 0.9,
 * beta2 0.999, eps 1e-8, decoupled weight decay on the weight matrices). */
itsca_status itsca_train_step(itsca_ctx* ctx, int first, int count, long mult, long offset, float pos_weight, float lr,
                              float weight_decay, float* loss, float* grad);
/* One epoch of AdamW steps of `batch` rows over a pseudo-shuffled order. */
itsca_status itsca_train_epoch(itsca_ctx* ctx, int batch, float lr, float weight_decay, float pos_weight,
                               uint64_t seed, float* mean_loss);
/* Logits of the first n training rows (the MFMA inference used by the filter). */
itsca_status itsca_train_scores(itsca_ctx* ctx, float* logits, int n);

/* Per-stage device time of the last run, from HIP events on the context's
 * stream (milliseconds): [0] cluster preparation and index tables,
 * [1] tracklets, [2] cells, [3] neighbours, [4] levels and roads.
 * Requires itsca_set_profiling(ctx, 1) before the run. */
itsca_status itsca_set_profiling(itsca_ctx* ctx, int enable);
itsca_status itsca_stage_times(itsca_ctx* ctx, float ms[5]);

/* Per-launch device time of the last profiled run (HIP events recorded on the
 * context's stream between consecutive launches); returns the number of
 * launches (filling at most cap entries), or -1 if no profiled run exists.
 * itsca_kernel_name gives the kernel of launch i. */
int itsca_kernel_times(itsca_ctx* ctx, float* ms, int cap);
const char* itsca_kernel_name(const itsca_ctx* ctx, int i);

#ifdef __cplusplus
}
#endif

#endif /* ITSCA_H */
