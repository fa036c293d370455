// cli_weight.cpp -- `multigpu_weight`, drop-in for the binary built from the
// reference's muca_delta_multigpu.cu (Makefile:54-55).
//
//   multigpu_weight L first_gpu last_gpu jacks starting_M recursive_update continuing
//                   [--walkers=W] [--seed=S] [--rounds=7|10] [--beta=B]
//                   [--max-iter=K] [--init-weight=PATH] [--outdir=DIR]
//
// Positional arguments, stdout lines and the files written under
// DIR/fss/L{L}/ follow muca_delta_multigpu.cu:87-546.  Differences, all
// deliberate:
//  - recursive_update=1 runs weight_recursion (:833-859, dead code in the
//    reference, whose argv[6] is parsed but unused); 0 runs the trivial update;
//  - each jack writes its own shortiter file (the reference keeps appending
//    to one stringstream, so jacks >= 1 get an unopenable concatenated name);
//  - a missing init-weight file is an error (the reference silently runs
//    from ln W = 0);
//  - --max-iter bounds the loop (the reference has no cap) and an
//    unconverged jack prints the reference's "Noooo...not converged";
//  - walkers default to the device's maxThreadsPerCU * CUs, as :100-106,
//    and any count is accepted.
// Everything numeric that the reference computes in 32-bit unsigned
// (total_run increments, Nwidth_step) keeps its 32-bit arithmetic.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "muca.h"

using namespace std;

namespace {

[[noreturn]] void die(const string &msg, int code = 1)
{
    cerr << "multigpu_weight: " << msg << endl;
    exit(code);
}

void check(int rc, const char *what)
{
    if (rc != MUCA_OK) die(string(what) + ": " + muca_last_error(), 2);
}

bool opt(const string &a, const char *name, string &val)
{
    const string p = string("--") + name + "=";
    if (a.compare(0, p.size(), p) == 0) {
        val = a.substr(p.size());
        return true;
    }
    return false;
}

void mkdirs(const string &path)
{
    string cur;
    stringstream ss(path);
    string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        mkdir(cur.c_str(), 0755);
    }
}

// jakknife (muca_delta_multigpu.cu:620-635)
void jakknife(const double *x, double *avg, double *err, unsigned J)
{
    check(muca_jackknife(x, J, avg, err), "jackknife");
}

}  // namespace

int main(int argc, char *argv[])
{
    if (argc < 8)
        die("usage: multigpu_weight L first_gpu last_gpu jacks starting_M recursive_update "
            "continuing [--walkers=W] [--seed=S] [--rounds=7|10] [--beta=B] [--max-iter=K] "
            "[--init-weight=PATH] [--outdir=DIR]");
    const unsigned L = (unsigned)atoi(argv[1]);
    const int first_gpu = atoi(argv[2]), last_gpu = atoi(argv[3]);
    const unsigned jacks = (unsigned)atoi(argv[4]);
    const unsigned initNUPDATES = (unsigned)atoi(argv[5]);
    const bool recursion = (bool)atoi(argv[6]);
    const bool extrapolation = (bool)atoi(argv[7]);
    if (last_gpu < first_gpu) die("last_gpu < first_gpu");
    if (jacks == 0) die("jacks must be >= 1");
    const unsigned NofGPU = 1 + last_gpu - first_gpu;

    uint32_t walkers = 0, seed = 100, rounds = 7;
    double beta = 2.0;
    unsigned long long max_iter = 0;
    string outdir = ".", init_path;
    for (int i = 8; i < argc; i++) {
        string a = argv[i], v;
        if (opt(a, "walkers", v)) walkers = (uint32_t)strtoul(v.c_str(), nullptr, 10);
        else if (opt(a, "seed", v)) seed = (uint32_t)strtoul(v.c_str(), nullptr, 10);
        else if (opt(a, "rounds", v)) rounds = (uint32_t)strtoul(v.c_str(), nullptr, 10);
        else if (opt(a, "beta", v)) beta = strtod(v.c_str(), nullptr);
        else if (opt(a, "max-iter", v)) max_iter = strtoull(v.c_str(), nullptr, 10);
        else if (opt(a, "init-weight", v)) init_path = v;
        else if (opt(a, "outdir", v)) outdir = v;
        else die("unknown option " + a);
    }
    const unsigned N = L * L;
    const string fss = outdir + "/fss/L" + argv[1];
    mkdirs(fss);

    vector<uint64_t> h_hist(N + 1), each(N + 1);
    vector<double> log_weight(N + 1, 0.0), p_acc(N + 1, 0.0);

    if (extrapolation) {  // :223-235
        if (init_path.empty()) init_path = fss + "/initweight_L128_T050.txt";
        ifstream f(init_path.c_str());
        if (!f) die("cannot open init weight file " + init_path);
        for (unsigned i = 0; i <= N; i++)
            if (!(f >> log_weight[i])) die("init weight file " + init_path + " has fewer than N+1 values");
    }

    // one context per GPU (muca_delta_multigpu.cu:93-205)
    vector<muca_ctx *> ctx(NofGPU, nullptr);
    for (unsigned l = 0; l < NofGPU; l++) {
        muca_params p{};
        p.L = L;
        p.walkers = walkers;
        p.device = first_gpu + (int)l;
        p.rounds = rounds;
        p.beta = beta;
        p.precision = MUCA_REF32;
        p.keying = MUCA_KEYING_WEIGHT;
        check(muca_create(&p, &ctx[l]), "muca_create");
    }
    muca_info info{};
    check(muca_query(ctx[0], &info), "muca_query");
    const unsigned NUM_WORKERS = info.walkers;


    unsigned width = 10;           // :314 (not reset between jacks)
    const unsigned therm_coeff = 1;
    const unsigned therm_run = therm_coeff * N;
    vector<double> total_kernel(jacks, 0.0), total_mem(jacks, 0.0), Nwidth_time(jacks, 0.0);
    vector<uint64_t> total_run(jacks, 0);
    vector<unsigned> Nwidth_step(jacks, 0);
    bool converged = false;

    for (unsigned mm = 0; mm < jacks; mm++) {
        total_run[mm] = 0;
        converged = false;
        bool reach = false;
        unsigned NUPDATES = initNUPDATES;
        for (unsigned l = 0; l < NofGPU; l++) {
            check(muca_init_lattice(ctx[l], mm, l, seed), "muca_init_lattice");   // :340-354
            check(muca_compute_state(ctx[l]), "muca_compute_state");               // :358-359
            vector<int32_t> D(NUM_WORKERS);
            check(muca_get_state(ctx[l], nullptr, D.data(), nullptr), "muca_get_state");
            cout << "E_DELTA : " << D[0] << endl;
            cout << "---------------------------------------- " << endl;
        }
        stringstream shortitrname;
        if (extrapolation)
            shortitrname << fss << "/shortiter_gpu" << NofGPU << "_T050c" << NUM_WORKERS << "_M"
                         << initNUPDATES << "_therm" << therm_coeff << "N_Jack" << mm
                         << "_extrapolated" << extrapolation << ".txt";
        else
            shortitrname << fss << "/shortiter_gpu" << NofGPU << "_T050c" << NUM_WORKERS << "_M"
                         << initNUPDATES << "_therm" << therm_coeff << "N_Jack" << mm
                         << "_iterupdate1.01_rand.txt";
        ofstream shortiter(shortitrname.str().c_str());
        shortiter << "#iter width D_KL mcs_time memcpy_time accumulated_time NUPDATES acc_NUP"
                  << endl;

        for (size_t k = 0;; k++) {  // :374
            if (max_iter && k >= max_iter) break;
            if (width == N) {
                NUPDATES = NUPDATES + 1;
                NUPDATES = (unsigned)(NUPDATES * 1.01);
            }
            total_run[mm] = total_run[mm] + (unsigned)(NUPDATES * NUM_WORKERS * NofGPU);
            for (unsigned l = 0; l < NofGPU; l++)
                check(muca_set_log_weight(ctx[l], log_weight.data(), N + 1), "set_log_weight");
            float kernel_time = 0, mem_time = 0;
            for (unsigned l = 0; l < NofGPU; l++) {
                float ms = 0;
                check(muca_run(ctx[l], (uint32_t)k, seed + 1000 * l, therm_run, NUPDATES,
                               nullptr, &ms),
                      "muca_run");
                kernel_time += ms;
            }
            std::fill(h_hist.begin(), h_hist.end(), 0);
            for (unsigned l = 0; l < NofGPU; l++) {
                float ms = 0;
                check(muca_get_hist(ctx[l], each.data(), N + 1, &ms), "muca_get_hist");
                mem_time += ms;
                for (unsigned i = 0; i <= N; i++) h_hist[i] += each[i];
            }
            total_kernel[mm] += (double)kernel_time;
            total_mem[mm] += (double)mem_time;

            int hist_start = 0, hist_end = 0;
            check(muca_hist_range(h_hist.data(), N, &hist_start, &hist_end), "hist_range");
            const unsigned width_temp = (unsigned)(hist_end - hist_start);
            if (width_temp > width) width = width_temp;
            const double D_KL = muca_kull_leibler(h_hist.data(), N);
            shortiter << k << " " << width_temp << " " << D_KL << " " << kernel_time * 1e-3 << " "
                      << mem_time * 1e-3 << " " << total_kernel[mm] * 1e-3 << " " << NUPDATES
                      << " " << total_run[mm] << endl;
            if ((width == N) && !reach) {
                cout << "Width reached N... Total mcs time : " << total_kernel[mm] * 1e-3 << endl;
                Nwidth_time[mm] = total_kernel[mm] * 1e-3;
                Nwidth_step[mm] = (unsigned)total_run[mm];
                reach = true;
            }
            if ((width == N) && (D_KL < 1e-4)) converged = true;
            if (converged) break;
            if (recursion)
                check(muca_update_recursive(h_hist.data(), N, hist_start, hist_end, p_acc.data(),
                                            log_weight.data()),
                      "update_recursive");
            else
                check(muca_update_trivial(h_hist.data(), N, log_weight.data()), "update_trivial");
            check(muca_normalise(N, log_weight.data()), "normalise");
        }
        shortiter.close();
        if (converged) {
            total_kernel[mm] = total_kernel[mm] * 1e-3;
            cout << "D_KL is below 1e-4 !! " << endl;
            cout << "Total run : " << total_run[mm] << "" << "\tN_width time : " << Nwidth_time[mm]
                 << "\tMCS time : " << total_kernel[mm] << "\tMEM time : " << total_mem[mm]
                 << "\tRatio of time : " << Nwidth_time[mm] / total_kernel[mm] << endl;
            stringstream weightname;
            if (extrapolation)
                weightname << fss << "/weight_gpu" << NofGPU << "_T050c" << NUM_WORKERS << "_M"
                           << initNUPDATES << "_therm" << therm_coeff << "N_JACK" << mm
                           << "_DKL1e-4_fromextrapolatedinit" << extrapolation << ".txt";
            else
                weightname << fss << "/weight_gpu" << NofGPU << "_T050c" << NUM_WORKERS << "_M"
                           << initNUPDATES << "_therm" << therm_coeff << "N_JACK" << mm
                           << "_DKL1e-4_iterupdate1.01_rand.txt";
            ofstream weightfile(weightname.str().c_str());
            for (unsigned i = 0; i <= N; i++)
                weightfile << i << " " << std::setprecision(10) << log_weight[i] << endl;
        } else {
            cout << "Noooo...not converged";
        }
    }

    // :498-524
    vector<double> total_run_d(jacks), Nwidth_step_d(jacks);
    for (unsigned i = 0; i < jacks; i++) {
        total_run_d[i] = (double)total_run[i];
        Nwidth_step_d[i] = (double)Nwidth_step[i];
    }
    double ttm = 0, tts = 0, tmm = 0, tms = 0, nwm = 0, nws = 0, trm = 0, trs = 0, nsm = 0, nss = 0;
    jakknife(total_kernel.data(), &ttm, &tts, jacks);
    jakknife(total_mem.data(), &tmm, &tms, jacks);
    jakknife(Nwidth_time.data(), &nwm, &nws, jacks);
    jakknife(total_run_d.data(), &trm, &trs, jacks);
    jakknife(Nwidth_step_d.data(), &nsm, &nss, jacks);
    if (converged) {
        stringstream pdname;
        if (extrapolation)
            pdname << fss << "/converge_result_gpu" << NofGPU << "_T050c" << NUM_WORKERS << "M"
                   << initNUPDATES << "_therm" << therm_coeff << "N_extrapolatedinit"
                   << extrapolation << ".txt";
        else
            pdname << fss << "/converge_result_gpu" << NofGPU << "_T050c" << NUM_WORKERS << "M"
                   << initNUPDATES << "_therm" << therm_coeff << "N_iterupdate1.01_rand.txt";
        ofstream pf(pdname.str().c_str());
        pf << "#total_time_mean" << " " << "total_time_std" << " " << "total_memtime_mean" << " "
           << "total_memtime_std" << " " << "Nwidth_time_mean" << " " << "Nwidth_time_std" << " "
           << "total_run_mean" << " " << "total_run_std" << " " << "Nwidth_step_mean" << " "
           << "Nwidth_step_std" << std::endl;
        pf << ttm << " " << tts << " " << " " << tmm << " " << tms << " " << nwm << " " << nws
           << " " << trm << " " << trs << " " << nsm << " " << nss << std::endl;
        pf << "#total_mcstime" << " " << "total_memtime" << " " << "Nwidth_time" << " "
           << "total_run" << " " << "Nwidth_step" << std::endl;
        for (unsigned i = 0; i < jacks; i++)
            pf << total_kernel[i] << " " << total_mem[i] << " " << Nwidth_time[i] << " "
               << total_run[i] << " " << Nwidth_step[i] << std::endl;
    }
    for (auto c : ctx) muca_destroy(c);
    return 0;
}
