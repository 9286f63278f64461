// Predict / TreeSHAP plans (placeholder until the serving kernels land).
#include <hip/hip_runtime.h>
#include "../../include/bonsai_hip.h"
extern "C" {
int bh_predictor_create(int, const bh_node *, const uint32_t *, uint32_t, const uint16_t *, uint32_t, float, float,
                        bh_predictor **out) { *out = nullptr; return 1; }
int bh_predictor_destroy(bh_predictor *) { return 0; }
int bh_predict(bh_predictor *, const uint8_t *, uint64_t, uint64_t, float *) { return 1; }
int bh_shap(bh_predictor *, const uint8_t *, uint64_t, uint64_t, double *) { return 1; }
int bh_predict_device(bh_predictor *, const uint8_t *, uint64_t, uint64_t, float *, void *) { return 1; }
int bh_shap_device(bh_predictor *, const uint8_t *, uint64_t, uint64_t, double *, void *) { return 1; }
}
