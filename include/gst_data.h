/* gst_data.h — synthetic graph generators and the reference's on-disk formats.
 *
 * CPU-only companion library (libgstdata.so) of the GST engine. It produces the
 * inputs the engine consumes: a simple undirected weighted graph restricted to its
 * largest connected component, stored as a symmetric CSR (beg_pos[n+1], adj[m], w[m],
 * m = number of directed arcs = 2 x undirected edges, adjacency sorted by neighbour id).
 *
 * File formats follow the reference readers:
 *   <name>.in            "V E" then "u v w" per undirected edge
 *                         (code/PrunedDP++/include/graph_hash_of_mixed_weighted_read_for_GSTP.h:15-66)
 *   <name>_beg_pos.bin   int64[V+1]  (code/D-TrimCDP-WB/include/graph.hpp:25,41)
 *   <name>_csr.bin       int64[m]
 *   <name>_weight.bin    int64[m]
 *   <name>.g             "g<n>:v v v", n 1-based, stored at index n-1 (read_for_GSTP.h:116-171)
 *   <name><k>.csv        k space-separated 0-based group ids per line (read_for_GSTP.h:172-218)
 *
 * Every function returns 0 / a non-NULL handle on success; on failure a negative code
 * (or NULL) and gstd_last_error() describes it.
 */
#ifndef GST_DATA_H
#define GST_DATA_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct gstd_graph gstd_graph; /* owned CSR graph */

/* R-MAT generator: 2^scale vertices, edge_factor*2^scale sampled edges with quadrant
 * probabilities (a,b,c,1-a-b-c); self loops dropped, duplicates merged, int32 weights
 * uniform in [wlo,whi] from a stream separate from the topology stream; restricted to
 * the largest connected component, relabelled 0..V'-1 in original-id order. */
gstd_graph *gstd_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                          int wlo, int whi, uint64_t seed);

/* Chung-Lu generator: n vertices with expected-degree weights (i+1)^(-1/(beta-1));
 * endpoints are drawn independently in proportion to the weights until m distinct
 * undirected non-loop edges exist; weights and LCC as above. */
gstd_graph *gstd_gen_chunglu(int64_t n, int64_t m, double beta, int wlo, int whi,
                             uint64_t seed);

/* Build from an undirected edge list (u[i], v[i], w[i]); the reader semantics of the
 * reference: both arcs are inserted and adjacency is sorted, duplicates are kept.
 * No LCC restriction. */
gstd_graph *gstd_from_edges(int64_t n, int64_t n_edges, const int32_t *u, const int32_t *v,
                            const int32_t *w);

int64_t gstd_num_vertices(const gstd_graph *g);
int64_t gstd_num_arcs(const gstd_graph *g);
const int64_t *gstd_beg_pos(const gstd_graph *g);
const int32_t *gstd_adj(const gstd_graph *g);
const int32_t *gstd_weight(const gstd_graph *g);
void gstd_free(gstd_graph *g);

/* writers / readers of the reference formats */
int gstd_write_in(const gstd_graph *g, const char *path);
int gstd_write_bin(const gstd_graph *g, const char *prefix);
gstd_graph *gstd_read_in(const char *path);
gstd_graph *gstd_read_bin(const char *prefix);

int gstd_write_groups(const char *path, int32_t n_groups, const int64_t *offs,
                      const int32_t *verts);
/* Reads a .g file; returns the number of groups, fills *offs (n+1) and *verts, which
 * the caller releases with gstd_free_buf. Members of each group are sorted ascending. */
int32_t gstd_read_groups(const char *path, int64_t **offs, int32_t **verts);
int gstd_write_queries(const char *path, int32_t n_queries, int32_t k, const int32_t *gids);
/* Reads <name><k>.csv; returns the number of rows, *k = entries in the first row. */
int32_t gstd_read_queries(const char *path, int32_t *k, int32_t **gids);
void gstd_free_buf(void *p);

const char *gstd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
