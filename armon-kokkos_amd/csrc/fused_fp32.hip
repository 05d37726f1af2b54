// Fused engine, FP32-mixed instantiation of the sweep kernels (sweep_kernels.hpp).
//
// Compiled with FMA contraction within each source expression and with the native-rate FP32
// reciprocal and square root: this form is held to its L1/Linf error against the FP64
// engine, not by bit parity; its results do not depend on the launch configuration (a cell's
// arithmetic is the same in every tile, run and unrolled copy of the march). The Bizarrium
// EOS, the CFL bound and the conservation sums stay in FP64.
#include "sweep_kernels.hpp"

namespace armon {

template void launch_sweep<float>(int, const SweepVariant&, const SweepArgs<float>&, hipStream_t);
template void launch_batch_sweep<float>(int, const SweepVariant&, const BatchArgs<float>&, const void*, int, int, int,
                                      int, hipStream_t);
template int batch_resident_blocks<float>(int, const SweepVariant&, int);

}  // namespace armon
