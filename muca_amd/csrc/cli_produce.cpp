// cli_produce.cpp -- `produce_gpu`, drop-in for the binary built from the
// reference's produce2_multigpu.cu (Makefile:42-43).
//
//   produce_gpu L first_gpu last_gpu range jacks therm_mult nup_mult delta0 gap
//               [--walkers=W] [--seed=S] [--rounds=7|10] [--jack-blocks=B]
//               [--weight-file=PATH] [--outdir=DIR] [--host-observables | --moments]
//
// Positional arguments, stdout and the observables file follow
// produce2_multigpu.cu:91-363: ln W is read from
// DIR/fss/L{L}/weight_mpi_T050c16_M200_JACK0_DKL1e-4.txt ("i lnW" lines), every
// Delta point runs a therm-only launch (iter 0, seed SEED+1000*gpu+l) and a
// production launch (iter l, seed SEED+2000*gpu) that fills the joint
// histogram H(D, M), and <|M|>, chi are reweighted at
// Delta = delta0 + gap*(l+1)/range with a jackknife over GPUs.
// Differences, all deliberate:
//  - reweighting runs in FP64 on the device from per-D moments of H(D,M)
//    (--host-observables keeps the reference's long-double host loop;
//    --moments accumulates exact per-D moments in the update kernel instead
//    of H(D,M), 6 (N+1) words per block in place of (N+1)(N+2)/2);
//  - --jack-blocks=B splits each GPU's walkers into B contiguous blocks and
//    jackknifes over all blocks (the reference's one-GPU jackknife is 0/0);
//  - a missing weight file is an error (the reference prints "!error!" and
//    then never leaves its read loop);
//  - histograms are zeroed before every launch (the reference zeroes them
//    inside the kernel, racing other workgroups), with 64-bit bin indices.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
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
    cerr << "produce_gpu: " << msg << endl;
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
    string cur, part;
    stringstream ss(path);
    if (!path.empty() && path[0] == '/') cur = "/";
    while (getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        mkdir(cur.c_str(), 0755);
    }
}

// observables (produce2_multigpu.cu:498-518), long double accumulation of
// double terms, 64-bit bin indices
void host_observables(double beta, unsigned N, const vector<uint64_t> &h,
                      const vector<double> &boltz, double *Mj, double *sj)
{
    long double sum = 0, M1 = 0, M2 = 0;
    for (uint64_t i = 0; i <= N; i++) {
        int64_t m;
        for (m = -(int64_t)i; m <= 0; m += 2) {
            const uint64_t c = h[(i * i + 2 * i + (uint64_t)m) / 2];
            sum = sum + (double)c * boltz[i];
            M1 = M1 + (double)((uint64_t)(-m) * c) * boltz[i];
            M2 = M2 + (double)((uint64_t)(m * m) * c) * boltz[i];
        }
        for (; m <= (int64_t)i; m += 2) {
            const uint64_t c = h[(i * i + 2 * i + (uint64_t)m) / 2];
            sum = sum + (double)c * boltz[i];
            M1 = M1 + (double)((uint64_t)m * c) * boltz[i];
            M2 = M2 + (double)((uint64_t)(m * m) * c) * boltz[i];
        }
    }
    *Mj = (double)(M1 / sum);
    *sj = (double)((M2 / sum - (M1 / sum) * (M1 / sum)) * beta / (long double)N);
}

}  // namespace

int main(int argc, char *argv[])
{
    if (argc < 10)
        die("usage: produce_gpu L first_gpu last_gpu range jacks therm_mult nup_mult delta0 gap "
            "[--walkers=W] [--seed=S] [--rounds=7|10] [--jack-blocks=B] [--weight-file=PATH] "
            "[--outdir=DIR] [--host-observables | --moments]");
    const unsigned L = (unsigned)stoi(argv[1]);
    const int first_gpu = stoi(argv[2]), last_gpu = stoi(argv[3]);
    if (last_gpu < first_gpu) die("last_gpu < first_gpu");
    const unsigned NofGPU = 1 + last_gpu - first_gpu;
    const unsigned range = (unsigned)stoi(argv[4]);
    if (range == 0) die("range must be >= 1");
    double delta = atof(argv[8]);
    const double beta = 1.0 / 0.5, delta_i = delta, gap = atof(argv[9]);

    uint32_t walkers = 0, seed = 100, rounds = 7, jack_blocks = 0;
    bool host_obs = false, moments = false;
    string outdir = ".", wpath;
    for (int i = 10; i < argc; i++) {
        string a = argv[i], v;
        if (opt(a, "walkers", v)) walkers = (uint32_t)strtoul(v.c_str(), nullptr, 10);
        else if (opt(a, "seed", v)) seed = (uint32_t)strtoul(v.c_str(), nullptr, 10);
        else if (opt(a, "rounds", v)) rounds = (uint32_t)strtoul(v.c_str(), nullptr, 10);
        else if (opt(a, "jack-blocks", v)) jack_blocks = (uint32_t)strtoul(v.c_str(), nullptr, 10);
        else if (opt(a, "weight-file", v)) wpath = v;
        else if (opt(a, "outdir", v)) outdir = v;
        else if (a == "--host-observables") host_obs = true;
        else if (a == "--moments") moments = true;
        else die("unknown option " + a);
    }
    if (host_obs && moments) die("--host-observables needs the joint histogram, not --moments");
    const unsigned N = L * L;
    const string fss = outdir + "/fss/L" + argv[1];
    mkdirs(fss);
    if (wpath.empty()) wpath = fss + "/weight_mpi_T050c16_M200_JACK0_DKL1e-4.txt";

    // weight file: "i lnW" per line (produce2_multigpu.cu:205-226)
    vector<double> log_weight(N + 1, 0.0);
    {
        ifstream in(wpath.c_str());
        if (!in) {
            cout << "!error!" << endl;
            die("cannot open weight file " + wpath);
        }
        for (unsigned i = 0; i <= N; i++) {
            int idx;
            if (!(in >> idx >> log_weight[i])) die("weight file " + wpath + " has fewer than N+1 lines");
        }
    }

    const uint32_t B = jack_blocks ? jack_blocks : 1;
    vector<muca_ctx *> ctx(NofGPU, nullptr);
    for (unsigned g = 0; g < NofGPU; g++) {
        muca_params p{};
        p.L = L;
        p.walkers = walkers;
        p.device = first_gpu + (int)g;
        p.rounds = rounds;
        p.beta = beta;
        p.precision = MUCA_FP64;
        p.keying = MUCA_KEYING_PRODUCE;
        p.joint_blocks = B;
        p.observables = moments ? MUCA_OBS_MOMENTS : MUCA_OBS_JOINT;
        check(muca_create(&p, &ctx[g]), "muca_create");
        check(muca_init_lattice(ctx[g], 0, g, seed), "muca_init_lattice");  // :126-140
        check(muca_set_log_weight(ctx[g], log_weight.data(), N + 1), "set_log_weight");
        check(muca_compute_state(ctx[g]), "muca_compute_state");            // :272-274
    }
    muca_info info{};
    check(muca_query(ctx[0], &info), "muca_query");
    const unsigned NUM_WORKERS = info.walkers;

    // unsigned 32-bit products as in :287,290
    const uint64_t therm_run = (unsigned)(N * (unsigned)atoi(argv[6]));
    const uint64_t NUPDATES = (unsigned)(N * (unsigned)atoi(argv[7]));
    cout << "\n TOTAL UPDATES : " << NUPDATES * NUM_WORKERS << " updates by each core : " << NUPDATES
         << " Therm run : " << therm_run << endl;

    const unsigned J = NofGPU * B;
    vector<double> avg_M(range), error_M(range), avg_sus(range), error_sus(range);
    vector<double> boltz(N + 1);
    const uint64_t jbins = (uint64_t)(N + 1) * (N + 2) / 2;
    vector<uint64_t> joint;
    for (unsigned l = 0; l < range; l++) {
        cout << "L" << argv[1] << " Progress...  " << 100 * l / range << "%" << "\r" << flush;
        delta = delta + gap / (double)range;
        // weight_shift (:391-400) and boltz (:297) on the host copy
        double mx = 0;
        for (unsigned i = 0; i < N + 1; i++) {
            const double v = -log_weight[i] - beta * delta * i;
            if (i == 0 || v > mx) mx = v;
        }
        for (unsigned i = 0; i < N + 1; i++) log_weight[i] = log_weight[i] + mx;
        for (unsigned j = 0; j < N + 1; j++) boltz[j] = exp(-beta * delta * j - log_weight[j]);

        for (unsigned g = 0; g < NofGPU; g++)
            check(muca_run(ctx[g], 0, seed + 1000 * g + l, therm_run, 0, nullptr, nullptr),
                  "muca_run(therm)");
        vector<double> Mj(J), sj(J);
        for (unsigned g = 0; g < NofGPU; g++) {
            check(muca_run(ctx[g], l, seed + 2000 * g, 0, NUPDATES, nullptr, nullptr),
                  "muca_run(production)");
            if (host_obs) {
                joint.resize(jbins);
                for (uint32_t b = 0; b < B; b++) {
                    check(muca_get_joint(ctx[g], b, joint.data(), jbins), "muca_get_joint");
                    host_observables(beta, N, joint, boltz, &Mj[g * B + b], &sj[g * B + b]);
                }
            } else {
                check(muca_reweight(ctx[g], &delta, 1, &Mj[g * B], &sj[g * B], nullptr),
                      "muca_reweight");
            }
        }
        check(muca_jackknife(Mj.data(), J, &avg_M[l], &error_M[l]), "jackknife");
        check(muca_jackknife(sj.data(), J, &avg_sus[l], &error_sus[l]), "jackknife");
    }

    stringstream pdname;
    pdname << fss << "/observables_gpu" << NofGPU << "_T050_JACKgpu_THERM" << argv[6] << "N_NUP"
           << argv[7] << "N_delta" << argv[8] << "to" << argv[9] << "noreset.txt";
    ofstream pf(pdname.str().c_str());
    for (size_t i = 0; i < range; i++)
        pf << delta_i + gap * (i + 1) / (double)range << " " << std::setprecision(10) << avg_M[i]
           << " " << std::setprecision(10) << error_M[i] << " " << std::setprecision(10)
           << avg_sus[i] << " " << std::setprecision(10) << error_sus[i] << std::endl;
    for (auto c : ctx) muca_destroy(c);
    return 0;
}
