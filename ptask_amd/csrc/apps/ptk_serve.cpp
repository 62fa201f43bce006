// ptk_serve.cpp -- a served stream of graph instances on the batched executor
// (include/ptk_batch.h): Poisson arrivals at a fixed rate, admission at iteration boundaries by
// HBM datablock footprint, continuous batching (instances that arrive while others run join the
// next iteration; finished ones leave and free their memory), p50/p99 latency beside throughput.
//
// Request mix (weights --mix F:G:B): FDTD cubic edge uniform on [32,192] with 20..100 steps,
// GMM N uniform on [10K, 1M], D = 16, K uniform on [8, 32], --gmm-steps EM steps, groupby rows
// uniform on [1M, 16M] (Zipf 1.1 over 65536 keys).  Latency = completion - scheduled arrival.
// --identity N re-runs N of the requests alone, each on a fresh executor, and compares their
// result checksums with the in-mix ones; --identity -1 re-runs every request alone (one at a
// time on one fresh executor, so no other instance is resident).  Prints one JSON object.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ptk_batch.h"
#include "ptk_prng.h"

namespace {

struct Options {
    std::string kernels = "ptask_amd/lib/kernels", mode = "batched";
    double qps = 20, duration = 10, arenaGb = 64, wF = 0.5, wG = 0.25, wB = 0.25;
    int requests = 0, identity = 8, gmmSteps = 20;
    int fdtdMin = 32, fdtdMax = 192, stepsMin = 20, stepsMax = 100, gmmNMin = 10000, gmmNMax = 1000000;
    long long gbMin = 1000000, gbMax = 16000000;
    unsigned long long seed = 0x5e7e;
    bool direct = false;  // batched iterations as direct launches instead of captured graphs
};

struct Req {
    int kind;
    long long a;
    int b, c, steps;
    unsigned long long seed;
    double arrival;  // ms after the start
    int id = -1;
    double done = -1;
};

int SubmitReq(ptk_batch *B, Req &r, int keep) {
    if (r.kind == PTK_INST_FDTD) return ptk_batch_submit_fdtd(B, (int)r.a, r.b, r.c, r.steps, r.seed, keep, &r.id);
    if (r.kind == PTK_INST_GROUPBY) return ptk_batch_submit_groupby(B, r.a, r.b, 1.1, r.seed, keep, &r.id);
    return ptk_batch_submit_gmm(B, (int)r.a, r.b, r.c, r.steps, r.seed, keep, &r.id);
}

double Pct(std::vector<double> v, double p) {
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    const double x = p * (v.size() - 1);
    const size_t i = (size_t)x;
    return i + 1 < v.size() ? v[i] + (x - i) * (v[i + 1] - v[i]) : v[i];
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
        else if (a == "--mode") o.mode = next();
        else if (a == "--direct") o.direct = true;
        else if (a == "--qps") o.qps = atof(next());
        else if (a == "--duration") o.duration = atof(next());
        else if (a == "--requests") o.requests = atoi(next());
        else if (a == "--arena-gb") o.arenaGb = atof(next());
        else if (a == "--identity") o.identity = atoi(next());
        else if (a == "--gmm-steps") o.gmmSteps = atoi(next());
        else if (a == "--fdtd-max") o.fdtdMax = atoi(next());
        else if (a == "--gb-max") o.gbMax = atoll(next());
        else if (a == "--gmm-n-max") o.gmmNMax = atoi(next());
        else if (a == "--seed") o.seed = strtoull(next(), nullptr, 0);
        else if (a == "--mix") {
            if (sscanf(next(), "%lf:%lf:%lf", &o.wF, &o.wG, &o.wB) != 3) { fprintf(stderr, "--mix F:G:B\n"); return 2; }
        } else { fprintf(stderr, "unknown option %s\n", a.c_str()); return 2; }
    }
    const int mode = o.mode == "plain" ? PTK_BATCH_PLAIN : o.mode == "batched" ? PTK_BATCH_BATCHED : -1;
    const double wsum = o.wF + o.wG + o.wB;
    if (mode < 0 || o.qps <= 0 || wsum <= 0 || o.fdtdMax < o.fdtdMin || o.gbMax < o.gbMin || o.gmmNMax < o.gmmNMin) {
        fprintf(stderr, "usage: ptk_serve [--mode batched|plain] [--qps Q] [--duration S | --requests N] "
                        "[--mix F:G:B] [--arena-gb G] [--identity N] [--gmm-steps S] [--seed S]\n");
        return 2;
    }
    // the request stream: exponential inter-arrival times (stream 40), kinds and sizes (41..46)
    std::vector<Req> reqs;
    double t = 0;
    for (uint64_t i = 0;; ++i) {
        if (o.requests > 0 && (int)i >= o.requests) break;
        t += -std::log(1.0 - ptk_rand_unit_f64(o.seed, 40, i)) * 1000.0 / o.qps;
        if (o.requests <= 0 && t > o.duration * 1000.0) break;
        Req r{};
        const double u = ptk_rand_unit_f64(o.seed, 41, i) * wsum;
        r.arrival = t;
        r.seed = o.seed ^ ((i + 1) * 0x9E3779B97F4A7C15ull);
        if (u < o.wF) {
            const int e = o.fdtdMin + (int)(ptk_rand_unit_f64(o.seed, 42, i) * (o.fdtdMax - o.fdtdMin + 1));
            r.kind = PTK_INST_FDTD;
            r.a = r.b = r.c = std::min(e, o.fdtdMax);
            r.steps = o.stepsMin + (int)(ptk_rand_unit_f64(o.seed, 43, i) * (o.stepsMax - o.stepsMin + 1));
            r.steps = std::min(r.steps, o.stepsMax);
        } else if (u < o.wF + o.wG) {
            r.kind = PTK_INST_GMM;
            r.a = o.gmmNMin + (long long)(ptk_rand_unit_f64(o.seed, 44, i) * (o.gmmNMax - o.gmmNMin + 1));
            r.a = std::min<long long>(r.a, o.gmmNMax);
            r.b = 16;
            r.c = std::min(32, 8 + (int)(ptk_rand_unit_f64(o.seed, 45, i) * 25));
            r.steps = o.gmmSteps;
        } else {
            r.kind = PTK_INST_GROUPBY;
            r.a = std::min(o.gbMax, o.gbMin + (long long)(ptk_rand_unit_f64(o.seed, 46, i) * (double)(o.gbMax - o.gbMin + 1)));
            r.b = 65535;
            r.steps = 1;
        }
        reqs.push_back(r);
    }
    if (reqs.empty()) {
        fprintf(stderr, "no requests in the window\n");
        return 2;
    }
    const unsigned long long arena = (unsigned long long)(o.arenaGb * 1073741824.0);
    ptk_batch *B = nullptr;
    if (ptk_batch_create(o.kernels.c_str(), arena, mode, &B)) {
        fprintf(stderr, "create: %s\n", ptk_batch_last_error());
        return 1;
    }
    if (o.direct) ptk_batch_set_launch(B, 0);
    // serve: submit what has arrived, run one iteration, repeat; idle until the next arrival
    const auto t0 = std::chrono::steady_clock::now();
    auto now = [&]() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); };
    size_t next = 0, completed = 0;
    int active = 0;
    double activeSum = 0;
    unsigned long long ticks = 0;
    while (completed < reqs.size()) {
        const double tn = now();
        while (next < reqs.size() && reqs[next].arrival <= tn) {
            if (SubmitReq(B, reqs[next], 0)) {
                fprintf(stderr, "submit: %s\n", ptk_batch_last_error());
                return 1;
            }
            ++next;
        }
        // run as many iterations back to back as fit before the next arrival (at least one)
        int maxIters = 1 << 30;
        if (next < reqs.size()) {
            unsigned long long st[7];
            ptk_batch_stats(B, st);
            const double perIter = st[0] ? st[6] / 1e3 / st[0] : 0.1;
            maxIters = (int)std::max(1.0, std::min(1e6, (reqs[next].arrival - now()) / std::max(perIter, 0.01)));
        }
        int done = 0, iters = 0;
        if (ptk_batch_run(B, -1, maxIters, &iters, &active, &done)) {
            fprintf(stderr, "run: %s\n", ptk_batch_last_error());
            return 1;
        }
        activeSum += (double)active * iters;
        ticks += (unsigned long long)iters;
        if (done) {
            completed = 0;
            for (Req &r : reqs) {
                if (r.id < 0) continue;
                int st;
                double ts, ta, td;
                ptk_batch_status(B, r.id, &st, &ts, &ta, &td);
                if (st >= PTK_STATE_DONE) {
                    if (r.done < 0) r.done = now();
                    ++completed;
                }
            }
        }
        if (active == 0 && next < reqs.size() && completed + (reqs.size() - next) == reqs.size()) {
            const double wait = reqs[next].arrival - now();
            if (wait > 0) std::this_thread::sleep_for(std::chrono::microseconds((long long)(wait * 1000)));
        }
    }
    const double wall = now();
    std::vector<double> lat, latK[3];
    double firstArr = reqs.front().arrival, lastDone = 0;
    for (Req &r : reqs) {
        const double l = r.done - r.arrival;
        lat.push_back(l);
        latK[r.kind].push_back(l);
        lastDone = std::max(lastDone, r.done);
    }
    // identity: some requests again, each alone on a fresh executor
    int idChecked = 0, idEqual = 0;
    if (o.identity < 0) {
        ptk_batch *A = nullptr;
        if (ptk_batch_create(o.kernels.c_str(), arena, mode, &A)) {
            fprintf(stderr, "create: %s\n", ptk_batch_last_error());
            return 1;
        }
        if (o.direct) ptk_batch_set_launch(A, 0);
        for (size_t i = 0; i < reqs.size(); ++i) {
            Req alone = reqs[i];
            unsigned long long a[8], m[8];
            const bool eq = !SubmitReq(A, alone, 0) && !ptk_batch_drain(A) && !ptk_batch_checksums(A, alone.id, a) &&
                            !ptk_batch_checksums(B, reqs[i].id, m) && !memcmp(a, m, sizeof(a));
            if (!eq) fprintf(stderr, "identity: request %zu (kind %d) differs alone and in the mix\n", i, reqs[i].kind);
            ++idChecked;
            idEqual += eq ? 1 : 0;
        }
        ptk_batch_destroy(A);
    } else if (o.identity > 0) {
        const size_t stride = std::max<size_t>(1, reqs.size() / (size_t)o.identity);
        for (size_t i = 0; i < reqs.size() && idChecked < o.identity; i += stride) {
            ptk_batch *A = nullptr;
            if (ptk_batch_create(o.kernels.c_str(), arena, mode, &A)) break;
            if (o.direct) ptk_batch_set_launch(A, 0);
            Req alone = reqs[i];
            unsigned long long a[8], m[8];
            bool eq = !SubmitReq(A, alone, 0) && !ptk_batch_drain(A) && !ptk_batch_checksums(A, alone.id, a) &&
                      !ptk_batch_checksums(B, reqs[i].id, m) && !memcmp(a, m, sizeof(a));
            ptk_batch_destroy(A);
            ++idChecked;
            idEqual += eq ? 1 : 0;
        }
    }
    unsigned long long st[7];
    ptk_batch_stats(B, st);
    size_t nk[3] = {0, 0, 0};
    for (Req &r : reqs) nk[r.kind]++;
    printf("{\"ok\": %s, \"mode\": \"%s%s\", \"qps\": %.3f, \"requests\": %zu, \"fdtd\": %zu, \"gmm\": %zu, \"groupby\": %zu, "
           "\"wall_ms\": %.1f, \"graphs_per_s\": %.3f, \"offered_per_s\": %.3f, \"p50_ms\": %.2f, \"p99_ms\": %.2f, "
           "\"p50_ms_by_kind\": [%.2f, %.2f, %.2f], \"p99_ms_by_kind\": [%.2f, %.2f, %.2f], \"ticks\": %llu, "
           "\"mean_active\": %.2f, \"launches\": %llu, \"graph_launches\": %llu, \"arena_peak\": %llu, "
           "\"arena_bytes\": %llu, \"identity_checked\": %d, \"identity_equal\": %d}\n",
           idEqual == idChecked ? "true" : "false", o.mode.c_str(), o.direct && mode == PTK_BATCH_BATCHED ? "-direct" : "", o.qps, reqs.size(), nk[0], nk[2], nk[1], wall,
           reqs.size() / ((lastDone - firstArr) / 1e3), reqs.size() / (reqs.back().arrival / 1e3), Pct(lat, 0.5),
           Pct(lat, 0.99), Pct(latK[0], 0.5), Pct(latK[2], 0.5), Pct(latK[1], 0.5), Pct(latK[0], 0.99),
           Pct(latK[2], 0.99), Pct(latK[1], 0.99), ticks, ticks ? activeSum / ticks : 0.0, st[1], st[2], st[4], st[5],
           idChecked, idEqual);
    ptk_batch_destroy(B);
    return idEqual == idChecked ? 0 : 1;
}
