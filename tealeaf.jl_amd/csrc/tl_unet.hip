// tl_unet.hip -- learned warm start (placeholder until the network lands).
#include "tl_internal.hpp"

namespace tl {
int unet_predict_guess(tl_ctx *c, double, double, int, double *)
{
    c->err = "learned warm start: no network loaded";
    return TL_E_STATE;
}
void unet_destroy(tl_ctx *) {}
int unet_record_step(tl_ctx *, double, double, int) { return TL_OK; }
}  // namespace tl

extern "C" int tl_unet_param_count(void) { return 0; }
extern "C" int tl_unet_set_weights(tl_ctx *, const float *) { return TL_E_STATE; }
extern "C" int tl_unet_get_weights(tl_ctx *, float *) { return TL_E_STATE; }
extern "C" int tl_unet_record(tl_ctx *, int) { return TL_E_STATE; }
extern "C" int tl_unet_dataset_size(tl_ctx *) { return 0; }
extern "C" int tl_unet_train(tl_ctx *, const tl_train_params *, float *) { return TL_E_STATE; }
extern "C" int tl_unet_predict(tl_ctx *, double, double, int) { return TL_E_STATE; }
