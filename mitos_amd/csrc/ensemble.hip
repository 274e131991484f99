// Ensembles: many realizations of uneven size issued as one graph launch.
//
// Realizations are grouped by (N, precision), and the realizations of a size
// up to 256^3 are dealt round-robin over three groups.  Each group owns one
// pipeline (its grids are reused realization after realization) and one
// stream; the groups are forked from a root stream and joined back, so groups
// overlap on the GPU while the realizations of one group follow each other.
// Pipelines of one size share their FFT schedule, so a realization gives the
// same bits whichever group runs it.  Every realization has its own spectrum table and its own
// P(k) slot, so the captured sequence needs no host work between
// realizations.  The FFT schedules of each group are autotuned first; the
// graph is captured once, after one eager pass has made every cached table
// (twiddles, windows, stencils, bin edges), and replayed by
// mitos_ensemble_run.
#include <algorithm>
#include <memory>

#include "pipe.h"

using namespace mitos;

struct EnsembleGroup {
    int N = 0;
    int precision = 0;
    Pipe<float>* pf = nullptr;
    Pipe<double>* pd = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::vector<int> members;
    int enqueue(const mitos_ic_config& c, const double* amp, double* pk, long long* obs) {
        return pf ? pf->enqueue(c, stream, false, amp, pk, obs)
                  : pd->enqueue(c, stream, false, amp, pk, obs);
    }
};

struct mitos_ensemble {
    std::vector<mitos_ic_config> cfg;
    std::vector<double> amplitude;
    std::vector<double*> amp;  // device phi1 table per realization
    std::vector<EnsembleGroup> groups;
    int bmax = 0;
    double* pk = nullptr;  // device [n][bmax][2]
    long long* obs = nullptr;  // device [n][bmax]
    hipStream_t root = nullptr;
    hipEvent_t fork = nullptr;
    // recorded on the launch stream right after each graph launch: the next
    // launch and the P(k) fetch wait for it
    hipEvent_t launched = nullptr;
    hipStream_t launch_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    // issue every realization on the group streams, forked from and joined to root
    int issue() {
        MITOS_HIP_CHECK(hipEventRecord(fork, root));
        for (auto& g : groups) {
            MITOS_HIP_CHECK(hipStreamWaitEvent(g.stream, fork, 0));
            for (int i : g.members)
                MITOS_TRY(g.enqueue(cfg[i], amp[i], pk + (size_t)i * 2 * bmax,
                                    obs + (size_t)i * bmax));
            MITOS_HIP_CHECK(hipEventRecord(g.done, g.stream));
            MITOS_HIP_CHECK(hipStreamWaitEvent(root, g.done, 0));
        }
        return MITOS_OK;
    }

    void release() {
        hipDeviceSynchronize();
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        for (auto& g : groups) {
            if (g.pf) {
                g.pf->destroy();
                delete g.pf;
            }
            if (g.pd) {
                g.pd->destroy();
                delete g.pd;
            }
            if (g.stream) hipStreamDestroy(g.stream);
            if (g.done) hipEventDestroy(g.done);
        }
        for (double* a : amp) hipFree(a);
        hipFree(pk);
        hipFree(obs);
        if (fork) hipEventDestroy(fork);
        if (launched) hipEventDestroy(launched);
        if (root) hipStreamDestroy(root);
    }
};

static int ensemble_build(mitos_ensemble* e) {
    const int n = (int)e->cfg.size();
    for (int i = 0; i < n; i++) {
        const mitos_ic_config& c = e->cfg[i];
        if (!size_supported(c.N) || c.bins < 1 || c.boxlen <= 0) return MITOS_ERR_ARG;
        if (c.M != 0 && c.M != c.N) return MITOS_ERR_UNSUPPORTED;  // lattice M = N
        if (c.precision != MITOS_FP32 && c.precision != MITOS_FP64) return MITOS_ERR_ARG;
        e->bmax = std::max(e->bmax, c.bins);
        auto it = std::find_if(e->groups.begin(), e->groups.end(), [&](const EnsembleGroup& g) {
            return g.N == c.N && g.precision == c.precision;
        });
        if (it == e->groups.end()) {
            EnsembleGroup g;
            g.N = c.N;
            g.precision = c.precision;
            e->groups.push_back(g);
            it = e->groups.end() - 1;
        }
        it->members.push_back(i);
    }
    {
        // the realizations of a size up to 256^3 dealt over three pipelines
        // (a 256^3 FP32 pipeline holds about 1.2 GB)
        std::vector<EnsembleGroup> split;
        for (auto& g : e->groups) {
            const int k = std::max(1, std::min((int)g.members.size(), g.N <= 256 ? 3 : 1));
            for (int j = 0; j < k; j++) {
                EnsembleGroup h;
                h.N = g.N;
                h.precision = g.precision;
                for (size_t m = j; m < g.members.size(); m += k) h.members.push_back(g.members[m]);
                split.push_back(h);
            }
        }
        e->groups = split;
    }
    MITOS_HIP_CHECK(hipStreamCreateWithFlags(&e->root, hipStreamNonBlocking));
    MITOS_HIP_CHECK(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
    MITOS_HIP_CHECK(hipEventCreateWithFlags(&e->launched, hipEventDisableTiming));
    for (auto& g : e->groups) {
        MITOS_HIP_CHECK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
        MITOS_HIP_CHECK(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
        if (g.precision == MITOS_FP32) {
            g.pf = new Pipe<float>();
            MITOS_TRY(g.pf->create(g.N));
            MITOS_TRY(g.pf->ensure_bins(e->bmax));
            MITOS_TRY(g.pf->autotune(g.stream));
        } else {
            g.pd = new Pipe<double>();
            MITOS_TRY(g.pd->create(g.N));
            MITOS_TRY(g.pd->ensure_bins(e->bmax));
            MITOS_TRY(g.pd->autotune(g.stream));
        }
    }
    e->amplitude.resize(n);
    e->amp.assign(n, nullptr);
    for (int i = 0; i < n; i++) {
        const mitos_ic_config& c = e->cfg[i];
        e->amplitude[i] = amplitude_cached(c.n_s, c.gamma, c.sigma8);
        const std::vector<double> h = phi1_host(c.N, c.boxlen, e->amplitude[i], c.n_s, c.gamma);
        MITOS_HIP_CHECK(hipMalloc(&e->amp[i], sizeof(double) * h.size()));
        MITOS_HIP_CHECK(copy_sync(e->amp[i], h.data(), sizeof(double) * h.size(),
                                  hipMemcpyHostToDevice));
    }
    MITOS_HIP_CHECK(hipMalloc(&e->pk, sizeof(double) * 2 * e->bmax * n));
    MITOS_HIP_CHECK(hipMalloc(&e->obs, sizeof(long long) * e->bmax * n));
    // eager pass: makes the cached tables and checks every launch
    MITOS_TRY(e->issue());
    MITOS_HIP_CHECK(hipStreamSynchronize(e->root));
    MITOS_HIP_CHECK(hipGetLastError());
    // capture
    MITOS_HIP_CHECK(hipStreamBeginCapture(e->root, hipStreamCaptureModeThreadLocal));
    const int rc = e->issue();
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(e->root, &graph);
    if (rc != MITOS_OK) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    MITOS_HIP_CHECK(ec);
    e->graph = graph;
    MITOS_HIP_CHECK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    return MITOS_OK;
}

extern "C" int mitos_ensemble_create(const struct mitos_ic_config* cfgs, int n,
                                     mitos_ensemble** out) {
    if (!cfgs || n < 1 || !out) return set_last_error(MITOS_ERR_ARG);
    pool_trim();
    mitos_ensemble* e = new mitos_ensemble();
    e->cfg.assign(cfgs, cfgs + n);
    const int rc = ensemble_build(e);
    if (rc != MITOS_OK) {
        e->release();
        delete e;
        return set_last_error(rc);
    }
    *out = e;
    return MITOS_OK;
}

extern "C" int mitos_ensemble_enqueue(mitos_ensemble* e, void* stream) {
    if (!e) return set_last_error(MITOS_ERR_ARG);
    hipStream_t s = stream ? (hipStream_t)stream : e->root;
    // the graph reuses one set of grids per group: a launch on another stream
    // waits for the previous one on the device
    if (e->launch_stream && e->launch_stream != s)
        MITOS_HIP_CHECK(hipStreamWaitEvent(s, e->launched, 0));
    MITOS_HIP_CHECK(hipGraphLaunch(e->exec, s));
    MITOS_HIP_CHECK(hipEventRecord(e->launched, s));
    e->launch_stream = s;
    return MITOS_OK;
}

extern "C" int mitos_ensemble_run(mitos_ensemble* e) {
    MITOS_TRY(mitos_ensemble_enqueue(e, nullptr));
    MITOS_HIP_CHECK(hipEventSynchronize(e->launched));
    MITOS_HIP_CHECK(hipGetLastError());
    return MITOS_OK;
}

extern "C" int mitos_ensemble_fetch_pk(mitos_ensemble* e, int i, struct mitos_ic_result* res) {
    if (!e || !res || i < 0 || i >= (int)e->cfg.size()) return set_last_error(MITOS_ERR_ARG);
    const int bins = e->cfg[i].bins;
    std::vector<double> sums(2 * bins);
    std::vector<long long> obs(bins);
    // the last launch, on whichever stream it was issued
    if (e->launch_stream) MITOS_HIP_CHECK(hipEventSynchronize(e->launched));
    MITOS_HIP_CHECK(hipStreamSynchronize(e->root));
    MITOS_HIP_CHECK(hipMemcpy(sums.data(), e->pk + (size_t)i * 2 * e->bmax,
                              sizeof(double) * 2 * bins, hipMemcpyDeviceToHost));
    MITOS_HIP_CHECK(hipMemcpy(obs.data(), e->obs + (size_t)i * e->bmax, sizeof(long long) * bins,
                              hipMemcpyDeviceToHost));
    std::vector<double> kk(bins), pp(bins);
    pk_finish(sums.data(), obs.data(), bins, e->cfg[i].boxlen, kk.data(), pp.data());
    if (res->k_in_bins) memcpy(res->k_in_bins, kk.data(), sizeof(double) * bins);
    if (res->power_in_bins) memcpy(res->power_in_bins, pp.data(), sizeof(double) * bins);
    if (res->obs_in_bins) memcpy(res->obs_in_bins, obs.data(), sizeof(long long) * bins);
    res->amplitude = e->amplitude[i];
    return MITOS_OK;
}

extern "C" int mitos_ensemble_destroy(mitos_ensemble* e) {
    if (!e) return MITOS_ERR_ARG;
    e->release();
    delete e;
    return MITOS_OK;
}
