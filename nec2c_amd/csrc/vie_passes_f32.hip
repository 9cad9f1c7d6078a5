// FP32 instantiation of the five-pass volume operator.
#include "vie_passes.h"

namespace vie {
template int volume_apply<float>(const VolumeView<float>&, const c32*, c32*, int, hipStream_t, hipEvent_t*,
                                   DevIndex);
template int pass_variant_count<float>(const Geom&, int);
template bool pass_variant_ok<float>(const Geom&, int, int);
template int pass_launch<float>(const VolumeView<float>&, int, int, const c32*, c32*, int, hipStream_t);
}  // namespace vie
