// Network transition queries on MI355X (gfx950).
#include "common.h"
#include "query.h"

namespace drw {

struct NetWeights {
    int kind = 0;
};

int net_weights_load(const char* models_txt, int kind, NetWeights** out)
{
    (void)models_txt;
    (void)kind;
    *out = nullptr;
    return set_error(-4, "network backend is not available in this build");
}

void net_weights_free(NetWeights* w) { delete w; }

int query_nn_launch(const NetWeights* w, const QueryArgs& q, uint8_t** scratch, size_t* cap, hipStream_t stream)
{
    (void)w; (void)q; (void)scratch; (void)cap; (void)stream;
    return set_error(-4, "network backend is not available in this build");
}

}  // namespace drw
