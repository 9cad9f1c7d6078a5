// FP32 instantiation of the five-pass volume operator.
#include "vie_passes.h"

namespace vie {
template int volume_apply<float>(const VolumeView<float>&, const c32*, c32*, int, hipStream_t, hipEvent_t*,
                                   DevIndex);
template int pass_variant_count<float>(const Geom&, int);
template bool pass_variant_ok<float>(const Geom&, int, int);
template unsigned pass_blocks_mj<float>(const Geom&, int, int);
template bool pass_mj_ok<float>(const Geom&, int, int);
template int volume_apply_mj<float>(const OpArgs<float>*, const unsigned*, const int*, const unsigned*,
                                  const MjClass*, int,
                                  hipStream_t);
template int pass_launch<float>(const VolumeView<float>&, int, int, const c32*, c32*, int, hipStream_t);
}  // namespace vie
