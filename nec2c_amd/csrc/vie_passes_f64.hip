// FP64 instantiation of the five-pass volume operator.
#include "vie_passes.h"

namespace vie {
template int volume_apply<double>(const VolumeView<double>&, const c64*, c64*, int, hipStream_t, hipEvent_t*,
                                   DevIndex);
template int pass_variant_count<double>(const Geom&, int);
template bool pass_variant_ok<double>(const Geom&, int, int);
template unsigned pass_blocks_mj<double>(const Geom&, int, int);
template bool pass_mj_ok<double>(const Geom&, int, int);
template int volume_apply_mj<double>(const OpArgs<double>*, const unsigned*, const int*, const unsigned*,
                                  const MjClass*, int,
                                  hipStream_t);
template int pass_launch<double>(const VolumeView<double>&, int, int, const c64*, c64*, int, hipStream_t);
}  // namespace vie
