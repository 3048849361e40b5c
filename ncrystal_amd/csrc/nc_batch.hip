// Many materials, one launch: a persistent kernel drains a queue of tiles
// (material, range of that material's neutrons).  Each material keeps its
// own composition view, RNG seed and call counter, and each of its neutrons
// draws from the same Philox stream as in a per-material call, so a batched
// launch returns exactly what one call per material returns.
//
// Tiles are handed out by one atomic counter in the order of a tile list
// built on the host (most expensive materials first, so that the tail of the
// launch is made of short tiles); the tile size and the number of resident
// workgroups are the autotuned parameters.
#include "nc_kernels.hh"

namespace ncamd {

// Mode 0: cross sections, mode 1: isotropic scatter sampling.
template <int Mode>
__global__ void __launch_bounds__(256) k_batch_persistent(const BatchMat* __restrict__ mats,
                                                          const BatchRng* __restrict__ rngs,
                                                          const BatchTile* __restrict__ tiles, int ntiles,
                                                          const double* __restrict__ ekin, double* __restrict__ o1,
                                                          double* __restrict__ o2, unsigned* __restrict__ next,
                                                          unsigned long long* __restrict__ nerr) {
  __shared__ int s_tile;
  while (true) {
    if (threadIdx.x == 0) s_tile = (int)atomicAdd(next, 1u);
    __syncthreads();
    const int t = s_tile;
    __syncthreads();
    if (t >= ntiles) break;  // every workgroup leaves once the list is drained
    const BatchTile tile = tiles[t];
    const BatchMat& m = mats[tile.mat];
    for (int64_t k = tile.begin + threadIdx.x; k < tile.end; k += blockDim.x) {
      const double E = ekin[m.offset + k];
      if (Mode == 0) {
        o1[m.offset + k] = comp_xs<false>(m.view, E, nullptr);
      } else {
        Rng rng;
        rng.init(rngs[tile.mat].seed, (uint32_t)rngs[tile.mat].stream, (uint64_t)k);
        double Ef, mu;
        const int rc = comp_sample_iso<kCompGeneral>(m.view, rng, E, Ef, mu);
        if (rc) { Ef = -1.0; mu = -999.0; atomicAdd(nerr, 1ull); }
        o1[m.offset + k] = Ef;
        o2[m.offset + k] = mu;
      }
    }
  }
}

__global__ void k_batch_reset(unsigned* next) { *next = 0u; }

void launch_batch(const BatchMat* mats, const BatchRng* rngs, const BatchTile* tiles, int ntiles, int nblocks,
                  int mode, const double* ekin, double* o1, double* o2, unsigned* next, unsigned long long* nerr,
                  hipStream_t st) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, st, next);
  if (mode == 0)
    hipLaunchKernelGGL(k_batch_persistent<0>, dim3(nblocks), dim3(256), 0, st, mats, rngs, tiles, ntiles, ekin, o1,
                       o2, next, nerr);
  else
    hipLaunchKernelGGL(k_batch_persistent<1>, dim3(nblocks), dim3(256), 0, st, mats, rngs, tiles, ntiles, ekin, o1,
                       o2, next, nerr);
}

}  // namespace ncamd
