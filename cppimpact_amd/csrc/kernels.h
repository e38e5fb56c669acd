// Launchers of the kernels in kernels.hip, called by the context code.
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace cpi {

size_t elem_block_smem(int npe, int prec, int phys, int ml, int max_el, int kmax);
int elem_block_threads(int npe, int rule);
int elem_block_default_elements(int npe);
int elem_direct_threads(int npe);

// part: 0 = the whole element pass; with the surrogate (PHYS_EP_MLP) 1
// launches only the elastic pass, 2 only the network on the queue and 3
// only the walk over the deferred blocks, so each can be timed on its own
hipError_t launch_elem_block(const Params &p, int rule, int prec, int phys, bool aux,
                             int nblocks, size_t smem, hipStream_t s, int part = 0);
hipError_t prepare_elem_block(int npe, int rule, int prec, int phys, size_t smem);
// Events the following kernel launches of this host thread record at the
// kernel's own start and end (hipExtLaunchKernel); (null, null) restores
// plain launches.  For per-kernel timing outside graph capture.
void set_launch_events(hipEvent_t start, hipEvent_t stop);
hipError_t launch_elem_direct(const Params &p, int rule, int prec, int phys, bool atomic,
                              hipStream_t s);
// prec: the element kernel's precision (its partials are stored in it)
hipError_t launch_node_bnd(const Params &p, int prec, bool aux, hipStream_t s);
hipError_t launch_node_all(const Params &p, bool aux, hipStream_t s);
hipError_t launch_node_init(const Params &p, hipStream_t s);
hipError_t launch_energy_reduce(const Params &p, hipStream_t s);
int energy_reduce_blocks();  // rows of p.ered (EP_N doubles each)
// lumped mass into p.m (scatter: 0 block, 1 colour range, 2 atomic)
hipError_t launch_mass(const Params &p, int rule, int scatter, int kind, double rho,
                       int nblocks, hipStream_t s);
hipError_t launch_elem_dt(const Params &p, double *out, hipStream_t s);
hipError_t launch_stress_elem(const Params &p, double *out, hipStream_t s);
hipError_t launch_debug_strain(const Params &p, int rule, double alpha, int dcase, double nu,
                               double *out, hipStream_t s);
// the surrogate network on queue rows [0, n) (rows hold ceil(n/16)*16 rows
// of 32 bytes; see mlp_infer.h)
hipError_t launch_mlp_rows(const MlpParams *P, uint16_t *rows, int n, hipStream_t s);
// copies n4 32-byte records
hipError_t launch_stream_copy(const void *src, void *dst, long n4, hipStream_t s);

}  // namespace cpi
