// Round trips recorded into hipGraphs: one launch call replays StP, the 2a->3a
// permute, c2r, r2c, the 3b->2b permute and PtS of one plan, or the grouped StP and
// PtS launches of a batch, on fixed device buffers.
#include <new>

#include "raysht_internal.h"

struct raysht_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

using namespace raysht;

extern "C" {

int raysht_graph_create_round_trip(raysht_graph **out, raysht_plan *p, int nphi, const double *s_in, double *work,
                                   double *p3, double *s_out)
{
    if (!out) return RAYSHT_EINVAL;
    *out = nullptr;
    if (!p || !s_in || !work || !p3 || !s_out || p->nr < 1) return RAYSHT_EINVAL;
    int rc = fft_prepare(nphi);
    if (rc != RAYSHT_OK) return rc;
    raysht_graph *g = new (std::nothrow) raysht_graph();
    if (!g) return RAYSHT_ENOMEM;
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        delete g;
        return RAYSHT_EHIP;
    }
    rc = RAYSHT_OK;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        rc = RAYSHT_EHIP;
    } else {
        void *sv = reinterpret_cast<void *>(st);
        rc = raysht_spectral_to_physical(p, nphi, s_in, work, p3, sv);
        if (rc == RAYSHT_OK) rc = raysht_physical_to_spectral(p, nphi, p3, work, s_out, sv);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc == RAYSHT_OK && e != hipSuccess) rc = RAYSHT_EHIP;
        g->graph = graph;
        if (rc == RAYSHT_OK && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
            rc = RAYSHT_EHIP;
    }
    (void)hipStreamDestroy(st);
    if (rc != RAYSHT_OK) {
        raysht_graph_destroy(g);
        return rc;
    }
    *out = g;
    return RAYSHT_OK;
}

int raysht_graph_create_batch(raysht_graph **out, raysht_batch *b, const double *const *s_in, double *const *p2,
                              double *const *s_out)
{
    if (!out) return RAYSHT_EINVAL;
    *out = nullptr;
    if (!b || !s_in || !p2 || !s_out) return RAYSHT_EINVAL;
    raysht_graph *g = new (std::nothrow) raysht_graph();
    if (!g) return RAYSHT_ENOMEM;
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        delete g;
        return RAYSHT_EHIP;
    }
    int rc = RAYSHT_OK;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        rc = RAYSHT_EHIP;
    } else {
        void *sv = reinterpret_cast<void *>(st);
        rc = raysht_batch_stp(b, s_in, p2, sv);
        if (rc == RAYSHT_OK) rc = raysht_batch_pts(b, p2, s_out, sv);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc == RAYSHT_OK && e != hipSuccess) rc = RAYSHT_EHIP;
        g->graph = graph;
        if (rc == RAYSHT_OK && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
            rc = RAYSHT_EHIP;
    }
    (void)hipStreamDestroy(st);
    if (rc != RAYSHT_OK) {
        raysht_graph_destroy(g);
        return rc;
    }
    *out = g;
    return RAYSHT_OK;
}

int raysht_graph_launch(raysht_graph *g, void *stream)
{
    if (!g || !g->exec) return RAYSHT_EINVAL;
    RAYSHT_HIP_TRY(hipGraphLaunch(g->exec, reinterpret_cast<hipStream_t>(stream)));
    return RAYSHT_OK;
}

void raysht_graph_destroy(raysht_graph *g)
{
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
}

}  // extern "C"
