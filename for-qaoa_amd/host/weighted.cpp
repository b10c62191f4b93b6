// weighted -- QAOA Max-Cut driver with the reference's command line:
//
//   weighted <P> <N> <D> <C> <B> [--graph FILE] [--angles FILE] [--dtype c128|c64]
//            [--schedule reference|fused|z2] [--dump FILE] [--report]
//
// Positionals and defaults (1 21 0 10 19) as src/Weighted.cpp:52-71.  Without
// --graph the instance is the reference's own: the complete graph K_N with unit
// weights (src/Weighted.cpp:86-88); without --angles every layer uses
// gamma = beta = pi/4 (src/Weighted.cpp:34-35).  The wall clock brackets the
// P-layer loop only (src/Weighted.cpp:113,140) and includes all device work.
// The first four output lines are the reference's (src/Weighted.cpp:153-159).
//
// Files: --graph holds one "u v w" edge per line (a first line "N" is
// optional and checked); --angles holds one "gamma beta" pair per layer.
// --dump writes the final amplitudes in LOGICAL qubit order as raw
// little-endian complex (the element type of --dtype), plus FILE.qmap with the
// physical->logical map of the state as it was left by the schedule.
//
// Any library error prints its message and exits with status 1, as the
// reference's checkCudaErrors does (include/helper_cuda.hpp:588-596).
#include "forqaoa.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

static void die(const char *what)
{
    fprintf(stderr, "weighted: %s: %s\n", what, fq_last_error());
    exit(1);
}

#define CHECK(call)             \
    do {                        \
        if ((call) != FQ_OK)    \
            die(#call);         \
    } while (0)

int main(int argc, char **argv)
{
    int P = 1, N = 21, D = 0, C = 10, B = 19;
    std::vector<std::string> pos;
    std::string graph_file, angles_file, dump_file, dtype_s = "c128", sched_s = "reference";
    bool report = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto need = [&](const char *name) -> std::string {
            if (i + 1 >= argc) {
                fprintf(stderr, "weighted: %s needs a value\n", name);
                exit(2);
            }
            return argv[++i];
        };
        if (a == "--graph")
            graph_file = need("--graph");
        else if (a == "--angles")
            angles_file = need("--angles");
        else if (a == "--dtype")
            dtype_s = need("--dtype");
        else if (a == "--schedule")
            sched_s = need("--schedule");
        else if (a == "--dump")
            dump_file = need("--dump");
        else if (a == "--report")
            report = true;
        else
            pos.push_back(a);
    }
    if (!pos.empty()) {
        if (pos.size() < 5) {
            fprintf(stderr, "usage: weighted <P> <N> <D> <C> <B> [options]\n");
            return 2;
        }
        P = atoi(pos[0].c_str());
        N = atoi(pos[1].c_str());
        D = atoi(pos[2].c_str());
        C = atoi(pos[3].c_str());
        B = atoi(pos[4].c_str());
    }
    const int dtype = dtype_s == "c64" ? FQ_C64 : FQ_C128;
    if (dtype_s != "c64" && dtype_s != "c128") {
        fprintf(stderr, "weighted: --dtype must be c128 or c64\n");
        return 2;
    }
    const int sched = sched_s == "fused" ? FQ_SCHED_FUSED : sched_s == "z2" ? FQ_SCHED_Z2 : FQ_SCHED_REFERENCE;
    if (sched_s != "fused" && sched_s != "reference" && sched_s != "z2") {
        fprintf(stderr, "weighted: --schedule must be reference, fused or z2\n");
        return 2;
    }

    fq_state *st = nullptr;
    CHECK(fq_state_create(N, D, B, C, dtype, &st));
    fq_graph *g = nullptr;
    CHECK(fq_graph_create(N, &g));
    if (graph_file.empty()) {
        for (int i = 0; i < N; i++)
            for (int j = i + 1; j < N; j++)
                CHECK(fq_graph_add_edge(g, i, j, 1.0));
    } else {
        std::ifstream f(graph_file);
        if (!f) {
            fprintf(stderr, "weighted: cannot open %s\n", graph_file.c_str());
            return 1;
        }
        std::string line;
        bool first = true;
        while (std::getline(f, line)) {
            if (line.empty() || line[0] == '#')
                continue;
            std::istringstream ss(line);
            std::vector<double> v;
            double x;
            while (ss >> x)
                v.push_back(x);
            if (first && v.size() == 1) {
                if ((int)v[0] != N) {
                    fprintf(stderr, "weighted: graph file is for N=%d, run has N=%d\n", (int)v[0], N);
                    return 1;
                }
                first = false;
                continue;
            }
            first = false;
            if (v.size() != 3) {
                fprintf(stderr, "weighted: bad edge line: %s\n", line.c_str());
                return 1;
            }
            CHECK(fq_graph_add_edge(g, (int)v[0], (int)v[1], v[2]));
        }
    }
    std::vector<double> gammas(P, M_PI_4), betas(P, M_PI_4);
    if (!angles_file.empty()) {
        std::ifstream f(angles_file);
        if (!f) {
            fprintf(stderr, "weighted: cannot open %s\n", angles_file.c_str());
            return 1;
        }
        for (int p = 0; p < P; ++p)
            if (!(f >> gammas[p] >> betas[p])) {
                fprintf(stderr, "weighted: %s has fewer than P=%d angle pairs\n", angles_file.c_str(), P);
                return 1;
            }
    }
    CHECK(fq_init_h(st));
    CHECK(fq_timing_enable(st, 1));

    const auto t0 = std::chrono::steady_clock::now();
    CHECK(fq_run(st, g, P, gammas.data(), betas.data(), sched));
    CHECK(fq_sync(st));
    const auto t1 = std::chrono::steady_clock::now();
    const double secs = std::chrono::duration<double>(t1 - t0).count();

    double rx_ms = 0, fused_ms = 0;
    CHECK(fq_timing_get(st, FQ_K_RX, &rx_ms, nullptr, nullptr));
    CHECK(fq_timing_get(st, FQ_K_FUSED, &fused_ms, nullptr, nullptr));
    printf("Total simulation time : %10.3f\n", secs);
    printf("Total RX time : %f ms\n", rx_ms + fused_ms);
    printf("Total csqs time : %f ms\n", 0.0);
    printf("Total transmission time : %f ms\n", 0.0);
    if (report) {
        double ex = 0, n2 = 0;
        CHECK(fq_expect_cut(st, g, &ex, &n2));
        uint64_t idx[8];
        double prob[8];
        const int k = N >= 3 ? 8 : (1 << N);
        CHECK(fq_topk(st, k, idx, prob));
        printf("Layers per second : %f\n", secs > 0 ? P / secs : 0.0);
        printf("Expected cut : %.15g\n", ex);
        printf("Norm squared : %.15g\n", n2);
        for (int i = 0; i < k; ++i)
            printf("Top %d : %llu %.15g\n", i, (unsigned long long)idx[i], prob[i]);
    }
    if (!dump_file.empty()) {
        size_t bytes = 0;
        CHECK(fq_state_info(st, nullptr, nullptr, nullptr, &bytes));
        std::vector<char> buf(bytes);
        CHECK(fq_state_copy_out(st, buf.data(), bytes, 1));
        FILE *f = fopen(dump_file.c_str(), "wb");
        if (!f || fwrite(buf.data(), 1, bytes, f) != bytes) {
            fprintf(stderr, "weighted: cannot write %s\n", dump_file.c_str());
            return 1;
        }
        fclose(f);
        std::vector<int> qmap(N);
        CHECK(fq_qubit_map(st, qmap.data()));
        FILE *m = fopen((dump_file + ".qmap").c_str(), "w");
        if (!m) {
            fprintf(stderr, "weighted: cannot write %s.qmap\n", dump_file.c_str());
            return 1;
        }
        for (int p = 0; p < N; ++p)
            fprintf(m, "%d%c", qmap[p], p + 1 < N ? ' ' : '\n');
        fclose(m);
    }
    fflush(stdout);
    fq_graph_destroy(g);
    fq_state_destroy(st);
    return 0;
}
