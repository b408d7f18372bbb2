// Batched stepping of many worlds (placeholder until the batch table lands).
#include "host_internal.h"

using namespace lenia;

extern "C" int lenia_batch_create(lenia_ctx *, lenia_world *const *, uint32_t, lenia_batch **out) {
    if (out) *out = nullptr;
    return fail(LENIA_ERR_UNSUPPORTED, "batches are not built yet");
}
extern "C" int lenia_batch_destroy(lenia_batch *) { return LENIA_OK; }
extern "C" int lenia_batch_step(lenia_batch *, uint32_t, uint32_t) {
    return fail(LENIA_ERR_UNSUPPORTED, "batches are not built yet");
}
