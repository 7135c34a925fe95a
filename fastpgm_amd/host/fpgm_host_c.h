/*
 * fpgm_host_c.h -- C entry points of the host mirror's structure comparison, for callers
 * that bind C (ctypes, cgo): the reference's CustomNetwork::LoadBIFFile,
 * Network::FindCompelled and BNSLComparison::GetSHD (CustomNetwork.cpp:52-230,
 * Network.cpp:547-684, BNSLComparison.cpp:12-73) over plain arrays.
 *
 * A learned graph is given as n_learned records (a, b, directed) in the order of the
 * reference's vec_edges: directed = 1 is the arc a -> b, directed = 0 the undirected
 * edge {a, b}.  A true DAG is given as (parent, child) pairs in vec_edges order.
 * Every function returns 0, or -1 with a message in fpgm_host_last_error().
 */
#ifndef FPGM_HOST_C_H
#define FPGM_HOST_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *fpgm_host_last_error(void);

/* SHD of a learned graph on n nodes against a true DAG (BNSLComparison::GetSHD). */
int32_t fpgm_host_shd(int32_t n, const int32_t *true_arcs, int32_t n_true, const int32_t *learned,
                      int32_t n_learned, int32_t *shd);

/* The same against the network of a BIF file (CustomNetwork::LoadBIFFile). */
int32_t fpgm_host_shd_bif(const char *bif_path, int32_t n, const int32_t *learned, int32_t n_learned,
                          int32_t *shd);

/* Structure of a BIF file: node count and (parent, child) arcs in the loader's order;
 * cards (cap_nodes entries, may be NULL) receives the domain sizes. */
int32_t fpgm_host_bif_arcs(const char *bif_path, int32_t *n_nodes, int32_t *cards, int32_t cap_nodes,
                           int32_t *arcs, int32_t cap_arcs, int32_t *n_arcs);

/* Chickering labels of a DAG's arcs (Network::OrderEdge + FindCompelled): labels[k] = 0
 * (compelled) or 1 (reversible) for arc k of the given order. */
int32_t fpgm_host_compelled(int32_t n, const int32_t *arcs, int32_t n_arcs, int32_t *labels);

#ifdef __cplusplus
}
#endif
#endif
