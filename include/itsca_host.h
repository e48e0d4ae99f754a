/* itsca_host.h -- C entry points of libitsca.so for event batches: the seeded
 * workload generator and the reference text formats (src/CA/IOUtils.cxx:38-123),
 * for callers that do not use the C++ API (Python, foreign-function bindings).
 */
#ifndef ITSCA_HOST_H
#define ITSCA_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct itsca_events itsca_events;

/* preset: 1 small (100 tracks), 2 central (8000 tracks), 3 mixed multiplicity
 * (log-uniform 50..8000), 5 pile-up (4 collisions); event i uses seed + i. */
itsca_events* itsca_events_generate(int preset, uint64_t seed, int num_events);
/* Reads the reference event format, or the binary side-car below (recognised
 * by its magic); NULL if the file cannot be opened or is malformed. */
itsca_events* itsca_events_load(const char* path);
/* Writes the reference event format (and the labels format if labels_path is
 * not NULL); returns 0 on success. */
int itsca_events_write(const itsca_events* ev, const char* events_path, const char* labels_path);
void itsca_events_free(itsca_events* ev);

/* Binary side-car (no counterpart in the reference): the same content as the
 * text format, little endian ("ITSCAEV1", event count; per event the vertex
 * count, 7 layer sizes, the vertices, then per layer x, y, z, alpha, cluster
 * id, mc), read back by itsca_events_load through mmap.  Returns 0 on success. */
int itsca_events_write_binary(const itsca_events* ev, const char* path);
/* Reads a labels file (IOUtils.cxx:80-123 format) into the batch, one list
 * per event (written back by itsca_events_write); returns 0 on success. */
int itsca_events_load_labels(itsca_events* ev, const char* labels_path);
/* Pile-up merger, the counterpart of generatePileUpData.cxx:44-207
 * (benchmarks/scripts/root): consecutive events are merged until a merged
 * event holds pile_up vertices; the MC labels of clusters and label rows that
 * follow the k-th vertex of a merged event are offset by (k - 1) * 100000.
 * Returns a new batch (free it with itsca_events_free), NULL on error. */
itsca_events* itsca_events_merge_pileup(const itsca_events* ev, int pile_up);

/* Road report of one event, as IOUtils::writeRoadsReport classifies it
 * (IOUtils.cxx:125-170), for a batch that holds labels (itsca_events_write
 * with labels, or itsca_events_load_labels): roads[v] holds num_roads[v] rows
 * of 8 int of vertex v (the ITSCA_ROADS rows of itsca.h: 5 cells, size,
 * label, fake).  Only labels of charged pions with 7 clusters count
 * (IOUtils.cxx:109).  counts receives the correct, duplicate and fake roads;
 * returns 0, or -1 if the batch has no labels for the event. */
int itsca_events_roads_report(const itsca_events* ev, int event, const int32_t* const* roads,
                              const int32_t* num_roads, int num_vertices, long counts[3]);

int itsca_events_count(const itsca_events* ev);
int itsca_events_num_vertices(const itsca_events* ev, int event);
/* offsets[8]: layer l occupies [offsets[l], offsets[l+1]) of the copy below. */
int itsca_events_layer_offsets(const itsca_events* ev, int event, int32_t offsets[8]);
/* Copies the event's clusters (all layers, layer after layer, file order) and
 * its vertices (3 floats each); returns 0 on success. */
int itsca_events_copy(const itsca_events* ev, int event, float* x, float* y, float* z, int32_t* cluster_id,
                      int32_t* mc_id, float* vertices);

/* Learned tracklet filter weights (no counterpart in the reference): reads a
 * weights file written by `python -m itsca.filter_training` (JSON: the 1441
 * parameters as float32 bit patterns in hex under "params_f32_bits", the logit
 * threshold under "threshold").  params receives n_params floats; returns 0 on
 * success, -1 if the file cannot be read or does not hold n_params values. */
int itsca_filter_load(const char* path, float* params, int n_params, float* threshold);
/* The phi gates of the same file ("phi_gates": one per tracklet layer, radians;
 * see itsca_set_filter_gates): gates receives n_gates floats, 0.3 each (no
 * gate) when the file has none; returns 0 on success, -1 if the file cannot be
 * read or its gate list is short. */
int itsca_filter_load_gates(const char* path, float* gates, int n_gates);

#ifdef __cplusplus
}
#endif

#endif /* ITSCA_HOST_H */
