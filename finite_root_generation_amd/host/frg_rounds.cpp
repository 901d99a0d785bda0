// frg_rounds.cpp -- the finite-root-generation round driver on top of the C ABI (include/frg.h).
//
// Host side of the reference's main (frg.cu:128-553) with the GPU work behind frg_track:
//   * resume: reads the root collection `fptColl` (roots of 21 complex doubles, 336 bytes each) and
//     the aggregate of error codes `errorCodesAgr` (int32 per trial) from the working directory if
//     they exist, recounts the totals and the number of rounds done (frg.cu:295-336);
//   * each round: tracks n paths (frg.cu:356-404), sorts, marks duplicates with code 6 and merges
//     the new roots into the collection (frg.cu:432-441, frgsupport.cu:10-215), updates the totals
//     and the Lambert-W completeness estimate (frg.cu:443-464), rewrites `fptColl` and appends the
//     round's codes to `errorCodesAgr` (frg.cu:472-480) and a statistics block to `progress`
//     (frg.cu:486-512).
// The three files keep the reference's byte and text formats, so a collection started by either
// program can be continued by the other.
//
// Path ids: trials are numbered consecutively across runs.  A round that starts after T trials
// (those recorded in `errorCodesAgr` included) tracks ids T .. T + n - 1 of the seed's Philox
// streams, so a resumed run continues with fresh paths whatever --paths the earlier runs used, and
// every result is reproducible.
//
//   frg_rounds NROUNDS [--paths N] [--seed S] [--fpm FILE] [--gamma RE,IM] [--dir DIR]
//                      [--persistent] [--mixed] [--learned] [--train-paths M] [--rounds-per-launch R]
//                      [--device D] [--tol T]
//
// --fpm reads the target as 30 complex doubles (480 bytes, the layout of Vect<nP>); the default is
// the reference's target (frg.cu:149-178) and gamma (frg.cu:41).  --learned trains the step-size
// model before the first round, on the step logs of M paths (default 32768) of the same target
// tracked from path ids 2^62 on (never used by a round), and tracks every round with it.
// --rounds-per-launch R tracks R rounds in one launch (frg_track_batch, one instance per round,
// same target) and then processes them one by one: the files are the same as with one launch per
// round, but a launch lasts at least as long as its longest path, so small rounds (the reference's
// 20,480 paths) leave the GPU mostly idle unless several share a launch.  The driver then holds R
// rounds of results in host memory (R x N x 340 bytes).
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "frg.h"

namespace {

// The reference's target parameters (frg.cu:149-178) and gamma (frg.cu:41).
const frg_c64 kRefFpm[] = {
    {4.241713408139979, -4.925334736425448},   {1.448005491494222, -2.686249989236069},
    {2.287911194543938, -0.9123108488455074},  {-1.870502483790705, -0.5778077414588303},
    {-0.7625329418400497, 0.9043788195368165}, {-4.142822753796651, 1.697671322829592},
    {1.023004534182935, -3.026210035710708},   {-2.522295755933046, 3.942421532145637},
    {-3.769870484797034, 0.7972377857916531},  {-4.285631161563732, -0.5650246890525263},
    {-2.166030800874488, 0.8836055774820508},  {4.579260949000442, -2.074013737196861},
    {4.801289733755713, -1.460721841824302},   {2.041266673362720, 2.962149232597863},
    {3.989999692714399, -0.2328295911150420},  {-3.996448356004510, 4.426854856711802},
    {2.315936719955769, -0.5527803857724418},  {-3.018911857947673, 0.9692384144638879},
    {-0.4906347357049512, 2.994316903874402},  {0.2397118512793472, -1.341440197750693},
    {4.942917423339324, -2.415216683074435},   {1.599536873563615, 2.367319121024760},
    {4.229853181495468, -1.149104287738972},   {-0.2180298061101595, 3.585004213517870},
    {4.311063203842579, -0.03375595047392820}, {4.938148066901551, 4.551701974206413},
    {-4.271214442569748, 0.1355349863802999},  {-0.7377434797048270, 1.831738837919296},
    {0.8894614668331116, 0.8657473695398501},  {4.036757863569754, 1.461164684969592},
};
static_assert(sizeof(kRefFpm) == sizeof(frg_c64) * FRG_NP, "30 target parameters");
const frg_c64 kRefGamma = {0.39231132891081666, 1.0062265257923928};

struct Args {
    int rounds = 0;
    int64_t paths = 20480;  // nPaths, frg.cu:25-27
    uint64_t seed = 1;
    std::string fpm_file, dir = ".";
    frg_c64 gamma = kRefGamma;
    bool persistent = false, mixed = false, learned = false;
    int64_t train_paths = 32768;
    int rounds_per_launch = 1;
    int device = 0;
    double tol = 1e-6;  // frgsupport.cu:7
};

[[noreturn]] void usage(const char *msg) {
    std::fprintf(stderr,
                 "%s\nusage: frg_rounds NROUNDS [--paths N] [--seed S] [--fpm FILE] [--gamma RE,IM] [--dir DIR]\n"
                 "                  [--persistent] [--mixed] [--learned] [--train-paths M] [--rounds-per-launch R]\n"
                 "                  [--device D] [--tol T]\n",
                 msg);
    std::exit(2);
}

Args parse(int argc, char **argv) {
    Args a;
    if (argc < 2 || std::sscanf(argv[1], "%i", &a.rounds) != 1 || a.rounds < 0) usage("nRound input error");
    for (int i = 2; i < argc; ++i) {
        const std::string k = argv[i];
        auto val = [&]() -> const char * {
            if (i + 1 >= argc) usage(("missing value for " + k).c_str());
            return argv[++i];
        };
        if (k == "--paths") a.paths = std::atoll(val());
        else if (k == "--seed") a.seed = std::strtoull(val(), nullptr, 10);
        else if (k == "--fpm") a.fpm_file = val();
        else if (k == "--dir") a.dir = val();
        else if (k == "--device") a.device = std::atoi(val());
        else if (k == "--tol") a.tol = std::atof(val());
        else if (k == "--gamma") {
            if (std::sscanf(val(), "%lf,%lf", &a.gamma.re, &a.gamma.im) != 2) usage("bad --gamma");
        } else if (k == "--persistent") a.persistent = true;
        else if (k == "--mixed") a.mixed = true;
        else if (k == "--learned") a.learned = true;
        else if (k == "--train-paths") a.train_paths = std::atoll(val());
        else if (k == "--rounds-per-launch") a.rounds_per_launch = std::atoi(val());
        else usage(("unknown option " + k).c_str());
    }
    if (a.paths <= 0) usage("--paths must be positive");
    if (a.train_paths <= 0) usage("--train-paths must be positive");
    if (a.rounds_per_launch <= 0) usage("--rounds-per-launch must be positive");
    return a;
}

std::string path_in(const Args &a, const char *name) { return a.dir + "/" + name; }

// Totals over error codes 0..6 (frg.cu:321-330, :444-452).
struct Tally {
    int64_t code[7] = {0, 0, 0, 0, 0, 0, 0};
    void add(const int32_t *e, int64_t n) {
        for (int64_t i = 0; i < n; ++i)
            if (e[i] >= 0 && e[i] <= 6) code[e[i]]++;
    }
    int64_t errors() const { return code[1] + code[2] + code[3] + code[4] + code[5]; }
};

void write_tally(std::ofstream &f, const char *title, const Tally &t, int64_t trials) {
    f << title << std::endl;
    f << "Roots collected:" << std::right << std::setw(10) << t.code[0] << std::endl;
    f << "Duplicates:     " << std::right << std::setw(10) << t.code[6] << std::endl;
    f << "Trials:         " << std::right << std::setw(10) << trials << std::endl;
    f << "Tracking errors:" << std::right << std::setw(10) << t.errors();
    f << "    (" << t.code[1] << " " << t.code[2] << " " << t.code[3] << " " << t.code[4] << " " << t.code[5] << ")"
      << std::endl;
    f << std::endl;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char **argv) {
    const Args a = parse(argc, argv);

    // target
    std::vector<frg_c64> fpm(kRefFpm, kRefFpm + FRG_NP);
    if (!a.fpm_file.empty()) {
        std::ifstream f(a.fpm_file, std::ios::binary);
        if (!f.read(reinterpret_cast<char *>(fpm.data()), sizeof(frg_c64) * FRG_NP)) {
            std::cerr << "cannot read 30 complex values from " << a.fpm_file << std::endl;
            return 1;
        }
    }

    // resume (frg.cu:295-336)
    std::vector<frg_c64> coll;
    {
        std::ifstream f(path_in(a, "fptColl"), std::ios::binary | std::ios::ate);
        if (f.is_open()) {
            const int64_t bytes = f.tellg();
            if (bytes % (int64_t)(sizeof(frg_c64) * FRG_NV) != 0) {
                std::cerr << "fptColl: size is not a multiple of one root" << std::endl;
                return 1;
            }
            coll.resize((size_t)(bytes / sizeof(frg_c64)));
            f.seekg(0);
            if (!f.read(reinterpret_cast<char *>(coll.data()), bytes)) {
                std::cerr << "fptColl: read failed" << std::endl;
                return 1;
            }
            std::cout << "Previous roots collected: " << coll.size() / FRG_NV << std::endl;
        }
    }
    Tally total;
    int64_t trials = 0, round = 0;
    {
        std::ifstream f(path_in(a, "errorCodesAgr"), std::ios::binary | std::ios::ate);
        if (f.is_open()) {
            const int64_t bytes = f.tellg();
            trials = bytes / (int64_t)sizeof(int32_t);
            std::vector<int32_t> codes((size_t)trials);
            f.seekg(0);
            if (!f.read(reinterpret_cast<char *>(codes.data()), trials * (int64_t)sizeof(int32_t))) {
                std::cerr << "errorCodesAgr: read failed" << std::endl;
                return 1;
            }
            total.add(codes.data(), trials);
            round = trials / a.paths;
            std::cout << "Previous trials: " << trials << std::endl;
            std::cout << "Previous rounds: " << round << std::endl;
        }
    }

    frg_opts o;
    frg_default_opts(&o);
    if (a.persistent) o.schedule = FRG_SCHED_PERSISTENT;
    if (a.mixed) o.precision = FRG_PREC_MIXED;
    frg_ctx *ctx = nullptr;
    int rc = frg_create(&ctx, a.device, &o);
    if (rc != FRG_OK) {
        std::cerr << "frg_create: " << frg_strerror(rc) << std::endl;
        return 1;
    }

    if (a.learned) {
        // step logs of the classical controller on the same target, then the model (GPU kernels)
        const auto t0 = std::chrono::steady_clock::now();
        const int64_t cap = std::min<int64_t>(a.train_paths * 128, INT32_MAX);
        std::vector<float> rec((size_t)cap * FRG_STEP_RECORD);
        int64_t nrec = 0;
        rc = frg_collect_step_log(ctx, fpm.data(), a.gamma, a.seed, (uint64_t)1 << 62, a.train_paths, rec.data(), cap,
                                  &nrec);
        std::vector<float> model(FRG_STEP_MODEL_SIZE);
        if (rc == FRG_OK && nrec == 0) rc = FRG_EINVAL;
        if (rc == FRG_OK) rc = frg_train_step_model(ctx, rec.data(), nrec, 4000, 4096, 2e-3f, 1, model.data(), nullptr);
        if (rc == FRG_OK) rc = frg_set_step_model(ctx, model.data());
        o.step_control = FRG_STEP_LEARNED;
        if (rc == FRG_OK) rc = frg_set_opts(ctx, &o);
        if (rc != FRG_OK) {
            std::cerr << "step model: " << frg_strerror(rc) << std::endl;
            frg_destroy(ctx);
            return 1;
        }
        std::cout << "Step model trained on " << nrec << " steps of " << a.train_paths << " paths in "
                  << seconds_since(t0) << " sec" << std::endl;
    }

    const int64_t n = a.paths;
    const int per_launch = std::min(a.rounds_per_launch, std::max(a.rounds, 1));
    std::vector<frg_c64> fpt_all((size_t)per_launch * n * FRG_NV), merged;
    std::vector<int32_t> err_all((size_t)per_launch * n);
    std::vector<frg_c64> fpms, gammas;
    for (int i = 0; i < per_launch; ++i) {
        fpms.insert(fpms.end(), fpm.begin(), fpm.end());
        gammas.push_back(a.gamma);
    }
    std::vector<uint64_t> seeds(per_launch, a.seed), first_ids(per_launch);
    std::vector<int64_t> counts(per_launch, n);
    double t_track = 0.0;
    std::cout << std::endl;
    for (int r = 1; r <= a.rounds; ++r) {
        const int64_t k = round + r - 1;  // rounds done before this one
        const int slot = (r - 1) % per_launch;
        if (slot == 0) {
            // track this round and the next ones of the launch: the i-th starts after trials + i*n
            // trials, so its path ids follow every id tracked before
            const int rr = std::min(per_launch, a.rounds - r + 1);
            for (int i = 0; i < rr; ++i) first_ids[i] = (uint64_t)(trials + i * n);
            frg_stats st;
            rc = rr == 1 ? frg_track(ctx, fpm.data(), a.gamma, a.seed, first_ids[0], n, fpt_all.data(), err_all.data(),
                                     nullptr, &st)
                         : frg_track_batch(ctx, rr, fpms.data(), gammas.data(), seeds.data(), first_ids.data(),
                                           counts.data(), fpt_all.data(), err_all.data(), nullptr, &st);
            if (rc != FRG_OK) {
                std::cerr << "!!! Homotopy GPU failure !!! " << frg_strerror(rc) << std::endl;
                frg_destroy(ctx);
                return 1;
            }
            t_track = st.total_ms * 1e-3 / rr;  // each round's share of the launch
        }
        const frg_c64 *fpt = fpt_all.data() + (size_t)slot * n * FRG_NV;
        int32_t *err = err_all.data() + (size_t)slot * n;
        std::cout << "Round " << k + 1 << std::endl;
        std::cout << n << " paths tracked in " << t_track << " sec" << std::endl << std::endl;

        // sort, delete duplicates, merge (frg.cu:432-441)
        auto t0 = std::chrono::steady_clock::now();
        const int64_t m = (int64_t)(coll.size() / FRG_NV);
        merged.resize((size_t)(m + n) * FRG_NV);
        const int64_t mm = frg_sort_dedupe_merge(coll.data(), m, fpt, err, n, a.tol, merged.data());
        if (mm < 0) {
            std::cerr << "frg_sort_dedupe_merge failed" << std::endl;
            frg_destroy(ctx);
            return 1;
        }
        merged.resize((size_t)mm * FRG_NV);
        coll.swap(merged);
        Tally rt;
        rt.add(err, n);
        for (int c = 0; c < 7; ++c) total.code[c] += rt.code[c];
        trials += n;
        const double ratio = (double)total.code[0] / ((double)total.code[0] + (double)total.code[6]);
        const double est = frg_completeness(total.code[0], total.code[6]);
        const double t_dedupe = seconds_since(t0);

        // files (frg.cu:472-480).  The collection is written to a temporary file that then
        // replaces fptColl, so a failed write (a full disk) leaves the previous collection whole;
        // either failure ends the run.
        t0 = std::chrono::steady_clock::now();
        {
            const std::string tmp = path_in(a, "fptColl.tmp");
            std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
            f.write(reinterpret_cast<const char *>(coll.data()), (std::streamsize)(coll.size() * sizeof(frg_c64)));
            f.close();
            if (!f || std::rename(tmp.c_str(), path_in(a, "fptColl").c_str()) != 0) {
                std::cerr << "cannot write fptColl in " << a.dir << std::endl;
                frg_destroy(ctx);
                return 1;
            }
        }
        {
            std::ofstream f(path_in(a, "errorCodesAgr"), std::ios::binary | std::ios::app);
            f.write(reinterpret_cast<const char *>(err), (std::streamsize)(n * sizeof(int32_t)));
            f.close();
            if (!f) {
                std::cerr << "cannot append to errorCodesAgr in " << a.dir << std::endl;
                frg_destroy(ctx);
                return 1;
            }
        }
        const double t_write = seconds_since(t0);

        // progress (frg.cu:486-512)
        std::ofstream p(path_in(a, "progress"), std::ios::app);
        p << "-----Progress statistics for round " << k + 1 << "-----" << std::endl;
        p << std::endl;
        write_tally(p, "---In Total---", total, trials);
        write_tally(p, "---Per Round---", rt, n);
        p << "---Progress---" << std::endl;
        p << "Success ratio:          " << std::right << std::setw(10) << ratio << std::endl;
        p << "Roots collected (est.): " << std::right << std::setw(10) << est * 100 << "%" << std::endl;
        p << std::endl;
        p << "--- Timing ---" << std::endl;
        p << "Tracking:                   " << std::right << std::setw(10) << t_track << " sec" << std::endl;
        p << "Sorting, del. dup., merging:" << std::right << std::setw(10) << t_dedupe << " sec" << std::endl;
        p << "Writing to disk:            " << std::right << std::setw(10) << t_write << " sec" << std::endl;
        p << std::endl << std::endl;
    }
    frg_destroy(ctx);
    return 0;
}
