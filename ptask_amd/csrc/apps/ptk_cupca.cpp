// ptk_cupca.cpp -- the reference's unrolled binary64 PCA graph (ptaskUnitTest/graphcupca.cpp
// PCAGraph::BuildGraph :245-1530 and RunTest :1551-1640) built against ptaskapi.h: one BLAS task
// per operation of pca_cpu's Gram-Schmidt loop, bound by name to ptaskDcopy, ptaskDaxpy,
// ptaskDgemv, ptaskDnrm2, ptaskDscal and ptaskDger with the reference's by-value parameter
// structs (graphcupca.cpp:161-220).  With --kernels pointing at pca.hsaco the tasks are the HIP
// kernels of kernels/pca_blas.h; at the oracle's host library, their host-task forms.
//
// Task sequence (graphcupca.cpp): without --prenorm the rows of R are centred first (Dcopy_1,
// Daxpy_1 for every other column, Daxpy_2 for every column); then per component k: Dcopy_2
// (T[:,k] = R[:,k]), J times {Dgemv_1 (P[:,k] = R^T T[:,k]); for k > 0 Dgemv_2 and Dgemv_3
// (Gram-Schmidt against P[:, :k]); Dnrm2_1 + Dscal_1 (normalise P[:,k]); Dgemv_4
// (T[:,k] = R P[:,k]); Dnrm2_2 + Dscal_2 (L[k] = |T[:,k]|, normalise)}, Dger_1
// (R -= L[k] T[:,k] P[:,k]^T); finally Dscal_3 (T[:,k] *= L[k]) per component.  Each buffer's
// first consumer takes its graph input channel; later consumers connect to its last producer.
//
// Differences from the reference graph, all of them corrections: Dgemv_4 passes M rows (the
// reference passes N, :1007, the same for its square test matrices); Dger_1's parameter block is
// fully initialised and uses the DGER template (the reference sets only OFFSETY and LDA, :1370,
// on a DSCAL-templated port of the same size); L is also a graph output.
//
// Prints one JSON object; --dump writes X (column-major), T, P, R and L as binary64.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ptaskapi.h"

using namespace PTask;

namespace {

// graphcupca.cpp:161-220
struct DCOPY_PARAMS { int N, OFFSETX, OFFSETY, INCX, INCY; };
struct DAXPY_PARAMS { int N; double ALPHA; int OFFSETX, OFFSETY, INCX, INCY; };
struct DGEMV_PARAMS { bool TRANSPOSE; int M, N; double ALPHA, BETA; int LDA, OFFSETX, OFFSETY, INCX, INCY; };
struct DNRM2_PARAMS { int N, OFFSETX, OFFSETY, INCX; };
struct DSCAL_PARAMS { int N; double ALPHA_CONST; int INCX, OFFSETX, OFFSETALPHA_VAR; bool INVERT_ALPHA_VAR; };
struct DGER_PARAMS { int OFFSETY, INCX, INCY, LDA, M, N, OFFSETALPHA_VAR, OFFSETX; };

struct Options {
    std::string kernels, dump, dist = "normal";
    int m = 512, n = 512, k = 3, j = 10, runs = 1;
    bool preNorm = false;
    unsigned long long seed = 7;
};

// One logical buffer of the graph (R, U, T, P, L): its graph input channel and its last producer.
struct Flow {
    const char *name;
    DatablockTemplate *tmpl = nullptr;
    GraphInputChannel *in = nullptr;
    Port *last = nullptr;
};

class Builder {
public:
    Builder(Graph *g, bool hip, int m, int n) : m_g(g), m_hip(hip), m_m(m), m_n(n) {}

    Port *In(Flow &f, UINT formal, UINT inout = (UINT)-1) {
        Port *p = Runtime::CreatePort(INPUT_PORT, f.tmpl, m_uid++, (char *)f.name, formal, inout);
        return p;
    }
    Port *Out(Flow &f, UINT formal) { return Runtime::CreatePort(OUTPUT_PORT, f.tmpl, m_uid++, (char *)f.name, formal); }
    // Connects a consumer port to the buffer's last producer (or makes it the graph input).
    void Feed(Flow &f, Port *p) {
        if (!f.last) {
            if (f.in) Runtime::HandleError("buffer %s has a second first consumer", f.name);
            f.in = m_g->AddInputChannel(p, (char *)(std::string(f.name) + "InputChannel").c_str());
        } else {
            m_g->AddInternalChannel(f.last, p, (char *)(std::string("InternalChannel_") + f.name).c_str());
        }
    }
    template <class P>
    Task *Add(CompiledKernel *k, const char *name, std::vector<Port *> ins, std::vector<Port *> outs,
              DatablockTemplate *tParams, const P &params) {
        Port *pp = Runtime::CreatePort(STICKY_PORT, tParams, m_uid++, (char *)"PARAMS", (UINT)ins.size());
        ins.push_back(pp);
        Task *t = m_g->AddTask(k, (UINT)ins.size(), ins.data(), (UINT)outs.size(), outs.data(), (char *)name);
        if (!t) return nullptr;
        if (m_hip) {
            t->SetComputeGeometry(m_m, m_n, 1);
            t->SetBlockAndGridSize(PTASKDIM3((m_m + 15) / 16, (m_n + 15) / 16, 1), PTASKDIM3(16, 16, 1));
        } else {
            t->SetComputeGeometry(1, 1, 1);  // a host task computes the whole operation per call
        }
        GraphInputChannel *c = m_g->AddInputChannel(pp, (char *)"paramsConstChannel");
        m_params.push_back({c, Runtime::AllocateDatablock(tParams, (void *)&params, sizeof(P), c)});
        ++m_tasks;
        return t;
    }
    void PushParams() {
        for (auto &pc : m_params) pc.first->Push(pc.second);
    }
    void ReleaseParams() {
        for (auto &pc : m_params) pc.second->Release();
        m_params.clear();
    }
    int Tasks() const { return m_tasks; }

private:
    Graph *m_g;
    bool m_hip;
    int m_m, m_n;
    UINT m_uid = 0;
    int m_tasks = 0;
    std::vector<std::pair<GraphInputChannel *, Datablock *>> m_params;
};

unsigned long long Fnv(const std::vector<double> &v) {
    unsigned long long h = 1469598103934665603ull;
    const unsigned char *b = (const unsigned char *)v.data();
    for (size_t i = 0; i < v.size() * sizeof(double); ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

}  // namespace

int main(int argc, char **argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) { fprintf(stderr, "missing value for %s\n", a.c_str()); exit(2); }
            return argv[++i];
        };
        if (a == "--kernels") o.kernels = next();
        else if (a == "--m") o.m = atoi(next());
        else if (a == "--n") o.n = atoi(next());
        else if (a == "--k") o.k = atoi(next());
        else if (a == "--j") o.j = atoi(next());
        else if (a == "--runs") o.runs = atoi(next());
        else if (a == "--prenorm") o.preNorm = true;
        else if (a == "--dist") o.dist = next();
        else if (a == "--seed") o.seed = strtoull(next(), nullptr, 0);
        else if (a == "--dump") o.dump = next();
        else o.kernels.clear(), i = argc;
    }
    if (o.kernels.empty() || o.m < 1 || o.n < 1 || o.k < 1 || o.k > o.n || o.j < 1 || o.runs < 1 ||
        (o.dist != "normal" && o.dist != "uniform")) {
        fprintf(stderr, "usage: ptk_cupca --kernels FILE [--m M] [--n N] [--k K<=N] [--j J] [--runs R] [--prenorm] "
                        "[--dist normal|uniform] [--seed S] [--dump file]\n");
        return 2;
    }
    const int M = o.m, N = o.n, K = o.k, J = o.j;
    Runtime::SetExitOnRuntimeFailure(FALSE);
    if (PTFAILED(Runtime::Initialize())) return 1;
    const bool hip = ptaskutils::SelectAcceleratorClass(o.kernels.c_str()) == ACCELERATOR_CLASS_HIP;
    auto kern = [&](const char *op) {
        CompiledKernel *k = Runtime::GetCompiledKernel((char *)o.kernels.c_str(), (char *)op);
        if (!k) {
            fprintf(stderr, "GetCompiledKernel %s failed: %s\n", op, Runtime::GetLastErrorText());
            exit(1);
        }
        return k;
    };
    CompiledKernel *kCopy = kern("ptaskDcopy"), *kAxpy = kern("ptaskDaxpy"), *kGemv = kern("ptaskDgemv");
    CompiledKernel *kNrm2 = kern("ptaskDnrm2"), *kScal = kern("ptaskDscal"), *kGer = kern("ptaskDger");
    Flow R{"R"}, U{"U"}, T{"T"}, P{"P"}, L{"L"};
    T.tmpl = Runtime::GetDatablockTemplate((char *)"T_double", sizeof(double), (UINT)M, (UINT)K, 1);
    P.tmpl = Runtime::GetDatablockTemplate((char *)"P_double", sizeof(double), (UINT)N, (UINT)K, 1);
    R.tmpl = Runtime::GetDatablockTemplate((char *)"R_double", sizeof(double), (UINT)M, (UINT)N, 1);
    U.tmpl = Runtime::GetDatablockTemplate((char *)"U_double", sizeof(double), (UINT)(M > N ? M : N), 1, 1);
    L.tmpl = Runtime::GetDatablockTemplate((char *)"L_double", sizeof(double), (UINT)K, 1, 1);
    DatablockTemplate *tCopy = Runtime::GetDatablockTemplate((char *)"DcopyParams", sizeof(DCOPY_PARAMS), PTPARM_BYVALSTRUCT);
    DatablockTemplate *tAxpy = Runtime::GetDatablockTemplate((char *)"DaxpyParams", sizeof(DAXPY_PARAMS), PTPARM_BYVALSTRUCT);
    DatablockTemplate *tGemv = Runtime::GetDatablockTemplate((char *)"DgemvParams", sizeof(DGEMV_PARAMS), PTPARM_BYVALSTRUCT);
    DatablockTemplate *tNrm2 = Runtime::GetDatablockTemplate((char *)"Dnrm2Params", sizeof(DNRM2_PARAMS), PTPARM_BYVALSTRUCT);
    DatablockTemplate *tScal = Runtime::GetDatablockTemplate((char *)"DscalParams", sizeof(DSCAL_PARAMS), PTPARM_BYVALSTRUCT);
    DatablockTemplate *tGer = Runtime::GetDatablockTemplate((char *)"DgerParams", sizeof(DGER_PARAMS), PTPARM_BYVALSTRUCT);

    Graph *g = new Graph();
    Builder b(g, hip, M, N);
    char name[96];
    bool ok = true;
    if (!o.preNorm) {
        {   // Dcopy_1: U = R[:,0]
            Port *r = b.In(R, 0, 0), *u = b.In(U, 1, 1);
            Port *ro = b.Out(R, 0), *uo = b.Out(U, 1);
            ok &= b.Add(kCopy, "Dcopy_1", {r, u}, {ro, uo}, tCopy, DCOPY_PARAMS{M, 0, 0, 1, 1}) != nullptr;
            b.Feed(R, r); b.Feed(U, u);
            R.last = ro; U.last = uo;
        }
        for (int c = 1; c < N; ++c) {  // Daxpy_1: U += R[:,c]
            Port *r = b.In(R, 0, 0), *u = b.In(U, 1, 1);
            Port *ro = b.Out(R, 0), *uo = b.Out(U, 1);
            snprintf(name, sizeof(name), "Daxpy_1_column_%d", c);
            ok &= b.Add(kAxpy, name, {r, u}, {ro, uo}, tAxpy, DAXPY_PARAMS{M, 1.0, c * M, 0, 1, 1}) != nullptr;
            b.Feed(R, r); b.Feed(U, u);
            R.last = ro; U.last = uo;
        }
        for (int c = 0; c < N; ++c) {  // Daxpy_2: R[:,c] += (-1/N) U
            Port *u = b.In(U, 0, 0), *r = b.In(R, 1, 1);
            Port *uo = b.Out(U, 0), *ro = b.Out(R, 1);
            snprintf(name, sizeof(name), "Daxpy_2_column_%d", c);
            ok &= b.Add(kAxpy, name, {u, r}, {uo, ro}, tAxpy, DAXPY_PARAMS{M, -1.0 / N, 0, c * M, 1, 1}) != nullptr;
            b.Feed(U, u); b.Feed(R, r);
            U.last = uo; R.last = ro;
        }
    }
    for (int k = 0; k < K; ++k) {
        {   // Dcopy_2: T[:,k] = R[:,k]
            Port *r = b.In(R, 0, 0), *t = b.In(T, 1, 1);
            Port *ro = b.Out(R, 0), *to = b.Out(T, 1);
            snprintf(name, sizeof(name), "Dcopy_2_component_%d", k);
            ok &= b.Add(kCopy, name, {r, t}, {ro, to}, tCopy, DCOPY_PARAMS{M, k * M, k * M, 1, 1}) != nullptr;
            b.Feed(R, r); b.Feed(T, t);
            R.last = ro; T.last = to;
        }
        for (int j = 0; j < J; ++j) {
            {   // Dgemv_1: P[:,k] = R^T T[:,k]
                Port *r = b.In(R, 0), *t = b.In(T, 1), *p = b.In(P, 2, 0), *po = b.Out(P, 0);
                snprintf(name, sizeof(name), "Dgemv_1_comp_%d_iter_%d", k, j);
                ok &= b.Add(kGemv, name, {r, t, p}, {po}, tGemv,
                            DGEMV_PARAMS{true, M, N, 1.0, 0.0, M, k * M, k * N, 1, 1}) != nullptr;
                b.Feed(R, r); b.Feed(T, t); b.Feed(P, p);
                P.last = po;
            }
            if (k > 0) {
                {   // Dgemv_2: U = P[:, :k]^T P[:,k]
                    Port *p = b.In(P, 0), *p2 = b.In(P, 1), *u = b.In(U, 2, 0), *uo = b.Out(U, 0);
                    snprintf(name, sizeof(name), "Dgemv_2_comp_%d_iter_%d", k, j);
                    ok &= b.Add(kGemv, name, {p, p2, u}, {uo}, tGemv,
                                DGEMV_PARAMS{true, N, k, 1.0, 0.0, N, k * N, 0, 1, 1}) != nullptr;
                    b.Feed(P, p); b.Feed(P, p2); b.Feed(U, u);
                    U.last = uo;
                }
                {   // Dgemv_3: P[:,k] -= P[:, :k] U
                    Port *p = b.In(P, 0), *u = b.In(U, 1), *p2 = b.In(P, 2, 0), *po = b.Out(P, 0);
                    snprintf(name, sizeof(name), "Dgemv_3_comp_%d_iter_%d", k, j);
                    ok &= b.Add(kGemv, name, {p, u, p2}, {po}, tGemv,
                                DGEMV_PARAMS{false, N, k, -1.0, 1.0, N, 0, k * N, 1, 1}) != nullptr;
                    b.Feed(P, p); b.Feed(U, u); b.Feed(P, p2);
                    P.last = po;
                }
            }
            {   // Dnrm2_1: L[k] = |P[:,k]|
                Port *p = b.In(P, 0), *l = b.In(L, 1, 0), *lo = b.Out(L, 0);
                snprintf(name, sizeof(name), "Dnrm2_1_comp_%d_iter_%d", k, j);
                ok &= b.Add(kNrm2, name, {p, l}, {lo}, tNrm2, DNRM2_PARAMS{N, k * N, k, 1}) != nullptr;
                b.Feed(P, p); b.Feed(L, l);
                L.last = lo;
            }
            {   // Dscal_1: P[:,k] /= L[k]
                Port *l = b.In(L, 0), *p = b.In(P, 1, 0), *po = b.Out(P, 0);
                snprintf(name, sizeof(name), "Dscal_1_comp_%d_iter_%d", k, j);
                ok &= b.Add(kScal, name, {l, p}, {po}, tScal, DSCAL_PARAMS{N, 1.0, 1, k * N, k, true}) != nullptr;
                b.Feed(L, l); b.Feed(P, p);
                P.last = po;
            }
            {   // Dgemv_4: T[:,k] = R P[:,k]
                Port *r = b.In(R, 0), *p = b.In(P, 1), *t = b.In(T, 2, 0), *to = b.Out(T, 0);
                snprintf(name, sizeof(name), "Dgemv_4_comp_%d_iter_%d", k, j);
                ok &= b.Add(kGemv, name, {r, p, t}, {to}, tGemv,
                            DGEMV_PARAMS{false, M, N, 1.0, 0.0, M, k * N, k * M, 1, 1}) != nullptr;
                b.Feed(R, r); b.Feed(P, p); b.Feed(T, t);
                T.last = to;
            }
            {   // Dnrm2_2: L[k] = |T[:,k]|
                Port *t = b.In(T, 0), *l = b.In(L, 1, 0), *lo = b.Out(L, 0);
                snprintf(name, sizeof(name), "Dnrm2_2_comp_%d_iter_%d", k, j);
                ok &= b.Add(kNrm2, name, {t, l}, {lo}, tNrm2, DNRM2_PARAMS{M, k * M, k, 1}) != nullptr;
                b.Feed(T, t); b.Feed(L, l);
                L.last = lo;
            }
            {   // Dscal_2: T[:,k] /= L[k]
                Port *l = b.In(L, 0), *t = b.In(T, 1, 0), *to = b.Out(T, 0);
                snprintf(name, sizeof(name), "Dscal_2_comp_%d_iter_%d", k, j);
                ok &= b.Add(kScal, name, {l, t}, {to}, tScal, DSCAL_PARAMS{M, 1.0, 1, k * M, k, true}) != nullptr;
                b.Feed(L, l); b.Feed(T, t);
                T.last = to;
            }
        }
        {   // Dger_1: R -= L[k] T[:,k] P[:,k]^T
            Port *l = b.In(L, 0), *t = b.In(T, 1), *p = b.In(P, 2), *r = b.In(R, 3, 0), *ro = b.Out(R, 0);
            snprintf(name, sizeof(name), "Dger_1_component_%d", k);
            ok &= b.Add(kGer, name, {l, t, p, r}, {ro}, tGer, DGER_PARAMS{k * N, 1, 1, M, M, N, k, k * M}) != nullptr;
            b.Feed(L, l); b.Feed(T, t); b.Feed(P, p); b.Feed(R, r);
            R.last = ro;
        }
    }
    for (int k = 0; k < K; ++k) {  // Dscal_3: T[:,k] *= L[k]
        Port *l = b.In(L, 0), *t = b.In(T, 1, 0), *to = b.Out(T, 0);
        snprintf(name, sizeof(name), "Dscal_3_component_%d", k);
        ok &= b.Add(kScal, name, {l, t}, {to}, tScal, DSCAL_PARAMS{M, 1.0, 1, k * M, k, false}) != nullptr;
        b.Feed(L, l); b.Feed(T, t);
        T.last = to;
    }
    if (!ok) {
        fprintf(stderr, "graph construction failed: %s\n", Runtime::GetLastErrorText());
        return 1;
    }
    GraphOutputChannel *oT = g->AddOutputChannel(T.last, (char *)"TOutputChannel");
    GraphOutputChannel *oP = g->AddOutputChannel(P.last, (char *)"POutputChannel");
    GraphOutputChannel *oR = g->AddOutputChannel(R.last, (char *)"ROutputChannel");
    GraphOutputChannel *oL = g->AddOutputChannel(L.last, (char *)"LOutputChannel");
    g->Run();
    b.PushParams();

    // X: M x N column-major from the shared seed (normal by Box-Muller, or uniform [0, 1))
    std::vector<double> X((size_t)M * N);
    unsigned long long s = o.seed * 0x9E3779B97F4A7C15ull + 1;
    auto u01 = [&]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return ((double)(s >> 11) + 0.5) / 9007199254740992.0;
    };
    for (size_t i = 0; i < X.size(); ++i)
        X[i] = o.dist == "uniform" ? u01() : std::sqrt(-2.0 * std::log(u01())) * std::cos(6.283185307179586 * u01());
    std::vector<double> zT((size_t)M * K, 0.0), zP((size_t)N * K, 0.0), zU((size_t)(M > N ? M : N), 0.0), zL((size_t)K, 0.0);
    std::vector<double> outT, outP, outR, outL;
    double ms = 0;
    for (int run = 0; run < o.runs; ++run) {
        auto t0 = std::chrono::steady_clock::now();
        auto push = [&](Flow &f, std::vector<double> &v) {
            if (!f.in) return;
            Datablock *blk = Runtime::AllocateDatablock(f.tmpl, v.data(), (UINT)(v.size() * sizeof(double)), f.in);
            f.in->Push(blk);
            blk->Release();
        };
        push(T, zT);
        push(P, zP);
        push(R, X);
        push(L, zL);
        push(U, zU);
        auto pull = [&](GraphOutputChannel *c, std::vector<double> &v, size_t n) {
            Datablock *blk = c->Pull(600000);
            if (!blk) return false;
            blk->Lock();
            const double *p = (const double *)blk->GetDataPointer(FALSE);
            v.assign(p, p + n);
            blk->Unlock();
            blk->Release();
            return true;
        };
        ok &= pull(oT, outT, (size_t)M * K) && pull(oP, outP, (size_t)N * K) && pull(oR, outR, (size_t)M * N) &&
              pull(oL, outL, (size_t)K);
        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    g->Stop();
    g->Teardown();
    b.ReleaseParams();
    Graph::DestroyGraph(g);
    Runtime::Terminate();
    if (ok && !o.dump.empty()) {
        FILE *f = fopen(o.dump.c_str(), "wb");
        if (!f) return 1;
        for (auto *v : {&X, &outT, &outP, &outR, &outL}) fwrite(v->data(), sizeof(double), v->size(), f);
        fclose(f);
    }
    std::string Ls;
    for (size_t i = 0; i < outL.size(); ++i) {
        char buf[40];
        snprintf(buf, sizeof(buf), "%s%.17g", i ? ", " : "", outL[i]);
        Ls += buf;
    }
    printf("{\"ok\": %s, \"backend\": \"%s\", \"m\": %d, \"n\": %d, \"k\": %d, \"j\": %d, \"prenorm\": %s, \"tasks\": %d, "
           "\"ms_last_run\": %.3f, \"L\": [%s], \"fnv1a_T\": \"%016llx\", \"fnv1a_R\": \"%016llx\"}\n",
           ok ? "true" : "false", hip ? "hip" : "host", M, N, K, J, o.preNorm ? "true" : "false", b.Tasks(), ms,
           Ls.c_str(), Fnv(outT), Fnv(outR));
    return ok ? 0 : 1;
}
