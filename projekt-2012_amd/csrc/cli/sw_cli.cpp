// SWalign command line: drop-in for Sterbic/Projekt-2012's
//     SWalign <seq1.fasta> <seq2.fasta> <match> <mismatch> <gap_open> <gap_extend>
// (README.md:51-57, src/main.cu:102-248).  The six positional arguments and the
// reference's stdout lines are kept; the device part runs through libswalign_hip.so.
// Optional flags may follow the six positional arguments:
//   --score-only           stop after the score and end cell
//   --checkpoint-rows K    checkpoint row spacing (default: chosen by the library)
//   --checkpoint-cols Kc   checkpoint column spacing
//   --no-cigar             do not print the alignment start and CIGAR lines
//   --visual               print the alignment in 50-column blocks
//   --gaps PREFIX          write PREFIX.first.gaps / PREFIX.second.gaps gap lists
//   --json                 print one JSON object with the result after the text
//   --host-checkpoints     keep checkpoint rows and columns in pinned host memory mapped
//                          into the device instead of HBM
//   --dump-rows DIR        write every checkpoint row as DIR/row_R in the reference's
//                          temp/row_R record format (src/Builders.cpp:27-104): the {int32 H,
//                          int32 F} pairs of the first R rows' last row, in column order
//   --device N             HIP device (default 0)
// Exit status: 255 for input errors, 254 for device errors (the reference's -1 / -2).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "swalign_hip.h"
#include "sw_fasta.h"
#include "sw_visual.h"

namespace {

const char* kVersion = "0.6.3";

[[noreturn]] void exit_with_msg(const char* msg, int code) {
    // src/SWutils.cpp:13-17
    printf("ERROR\n%s\n\n", msg);
    fflush(stdout);
    exit(code);
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char** argv) {
    const auto t_start = std::chrono::steady_clock::now();
    printf("### Welcome to SWalign v%s\n\n", kVersion);

    std::vector<const char*> pos;
    bool score_only = false, no_cigar = false, visual = false, json = false, host_ck = false;
    const char* dump_dir = nullptr;
    int64_t K = 0, Kc = 0;
    int device = 0;
    const char* gaps_prefix = nullptr;
    for (int k = 1; k < argc; ++k) {
        const char* a = argv[k];
        if (pos.size() >= 6 && strncmp(a, "--", 2) == 0) {
            if (!strcmp(a, "--score-only")) score_only = true;
            else if (!strcmp(a, "--no-cigar")) no_cigar = true;
            else if (!strcmp(a, "--visual")) visual = true;
            else if (!strcmp(a, "--json")) json = true;
            else if (!strcmp(a, "--host-checkpoints")) host_ck = true;
            else if (!strcmp(a, "--dump-rows") && k + 1 < argc) dump_dir = argv[++k];
            else if (!strcmp(a, "--checkpoint-rows") && k + 1 < argc) K = atoll(argv[++k]);
            else if (!strcmp(a, "--checkpoint-cols") && k + 1 < argc) Kc = atoll(argv[++k]);
            else if (!strcmp(a, "--gaps") && k + 1 < argc) gaps_prefix = argv[++k];
            else if (!strcmp(a, "--device") && k + 1 < argc) device = atoi(argv[++k]);
            else {
                printf("Unknown option %s\n\n", a);
                return -1;
            }
        } else {
            pos.push_back(a);
        }
    }
    if (pos.size() != 6) {
        // src/main.cu:107-110
        printf("Expected 6 input arguments, not %d!\n\n", (int)pos.size());
        return -1;
    }

    FastaRecord first, second;
    printf("> Loading input sequences... ");
    if (!fasta_load(pos[0], &first) || !fasta_load(pos[1], &second))
        exit_with_msg("An error has occured while loading input sequences.", -1);
    printf("DONE\n\n");
    printf("First sequence of length %lld:\n%s\n\n", (long long)first.seq.size(), first.name.c_str());
    printf("Second sequence of length %lld:\n%s\n\n", (long long)second.seq.size(), second.name.c_str());

    printf("> Looking for HIP capable devices... ");
    fflush(stdout);
    swh_ctx* ctx = nullptr;
    int rc = swh_ctx_create(device, &ctx);
    if (rc == SWH_ERR_NO_DEVICE) exit_with_msg("No HIP capable device detected.", -2);
    if (rc != SWH_OK) exit_with_msg(swh_strerror(rc), -2);
    char name[256];
    int32_t cus = 0, clk = 0;
    int64_t hbm = 0;
    swh_device_info(ctx, name, sizeof name, &cus, &clk, &hbm);
    printf("DONE\n\n");
    printf("Found HIP capable GPU, picked GPU number %d:\n", device + 1);
    printf("\t>Name: %s\n", name);
    printf("\t>Global memory: %.0f MBytes\n", (double)hbm / 1048576.0);
    printf("\t>Compute units: %d\n", cus);
    printf("\t>Stream processors: %d\n\n", cus * 64);

    printf("> Initializing launch configuration... DONE\n\n");
    printf("Launch configuration:\n");
    printf("\t>Wavefront: 64 lanes, %d rows per lane\n", 4);
    printf("\t>Stripe height: %d rows\n\n", 256);

    printf("> Preparing SWquery... DONE\n\n");

    // src/SWutils.cpp:26-47 (validated after device set-up, as in src/main.cu:145)
    printf("> Initializing scoring values... ");
    swh_scoring sc{atoi(pos[2]), atoi(pos[3]), atoi(pos[4]), atoi(pos[5])};
    if (sc.match < 1 || sc.mismatch > -1 || sc.gap_open > -1 || sc.gap_extend > -1)
        exit_with_msg("One or more scoring values were not usable!", -1);
    printf("DONE\n\nScoring values:\n");
    printf("\t>Match: %d\n", sc.match);
    printf("\t>Mismatch: %d\n", sc.mismatch);
    printf("\t>First gap: %d\n", sc.gap_open);
    printf("\t>Gap extension: %d\n\n", sc.gap_extend);

    printf("> Starting alignment process... ");
    fflush(stdout);
    swh_opts opts{};
    opts.checkpoint_rows = K;
    opts.checkpoint_cols = Kc;
    opts.score_only = score_only ? 1 : 0;
    opts.want_xpoints = 1;
    opts.host_checkpoints = host_ck ? 1 : 0;
    swh_result res{};
    const double t_before = seconds_since(t_start);
    rc = swh_align_pair(ctx, first.seq.data(), (int64_t)first.seq.size(), second.seq.data(),
                        (int64_t)second.seq.size(), &sc, &opts, &res);
    const double t_after = seconds_since(t_start);
    if (rc != SWH_OK) exit_with_msg(swh_strerror(rc), rc == SWH_ERR_SCORING || rc == SWH_ERR_INVALID_ARG ? -1 : -2);
    printf("DONE\n\n");

    const double app_s = t_before + res.ms_forward / 1000.0;
    const double cells = (double)first.seq.size() * (double)second.seq.size();
    // src/main.cu:220-224: GCUPS over the application time up to the end of stage 1
    printf("\t>Kernel executed in %f s\n", res.ms_forward / 1000.0);
    printf("\t>Application executed in %f s\n", app_s);
    printf("\t>Cell updates per second: %lf GCUPS\n", cells / 1e6 / (app_s * 1000.0));
    const int64_t er = res.score > 0 ? res.end_row + 1 : 1, ec = res.score > 0 ? res.end_col + 1 : 1;
    printf("\t>Alignment score: %d at [%lld, %lld]\n\n", res.score, (long long)er, (long long)ec);

    const uint8_t* A = res.swapped ? second.seq.data() : first.seq.data();
    const uint8_t* B = res.swapped ? first.seq.data() : second.seq.data();
    if (!score_only) {
        printf("> Creating Crosspointer object... DONE\n\n");
        printf("> Starting crosspointing process... DONE\n\n");
        printf("Found %lld crosspoints:\n", (long long)res.n_xpoints);
        for (int64_t k = 0; k < res.n_xpoints; ++k)
            printf("\t>xPoint[%lld]: [%lld, %lld] = %d\n", (long long)k, (long long)res.xpoints[k].row,
                   (long long)res.xpoints[k].col, res.xpoints[k].score);
        printf("\n");
        if (!no_cigar && res.score > 0) {
            printf("\t>Alignment start: [%lld, %lld]\n", (long long)res.start_row + 1, (long long)res.start_col + 1);
            printf("\t>Alignment end: [%lld, %lld]\n", (long long)er, (long long)ec);
            printf("\t>CIGAR: %s\n", res.cigar ? res.cigar : "");
            printf("\t>Stage times: forward %.3f ms, end search %.3f ms, traceback %.3f ms\n",
                   res.ms_forward, res.ms_end_search, res.ms_traceback);
            printf("\t>Full alignment: %f s, %lf GCUPS\n\n", (t_after - t_before),
                   cells / 1e9 / std::max(t_after - t_before, 1e-9));
        }
        if (visual && res.score > 0 && res.cigar)
            print_alignment(stdout, A, B, res.start_row, res.start_col, res.cigar);
        if (gaps_prefix && res.score > 0 && res.cigar) {
            std::vector<GapRun> ga, gb;
            cigar_gap_lists(res.cigar, res.start_row, res.start_col, &ga, &gb);
            const std::string p1 = std::string(gaps_prefix) + ".first.gaps";
            const std::string p2 = std::string(gaps_prefix) + ".second.gaps";
            if (!write_gap_list(p1.c_str(), ga) || !write_gap_list(p2.c_str(), gb))
                exit_with_msg("Error writing gap lists.", -1);
        }
    }
    if (dump_dir && !score_only && res.checkpoint_rows > 0 && res.m / res.checkpoint_rows > 0) {
        // the checkpoint rows of the same spacing, recomputed by a forward pass and written as
        // the reference's stage 1 left them in temp/ (file row_R: matrix row R-1, n records)
        const int64_t K = res.checkpoint_rows, Kc = res.checkpoint_cols > 0 ? res.checkpoint_cols : K;
        const int64_t nrows = res.m / K, n = res.n;
        std::vector<int32_t> rows((size_t)(nrows * n * 2));
        int32_t score = 0;
        int64_t i = 0, j = 0;
        rc = swh_forward(ctx, A, res.m, B, n, &sc, K, Kc, &score, &i, &j, nullptr, rows.data(), nullptr);
        if (rc != SWH_OK) exit_with_msg(swh_strerror(rc), -2);
        for (int64_t q = 0; q < nrows; ++q) {
            const std::string path = std::string(dump_dir) + "/row_" + std::to_string((q + 1) * K);
            FILE* f = fopen(path.c_str(), "wb");
            if (!f || fwrite(rows.data() + q * n * 2, sizeof(int32_t) * 2, (size_t)n, f) != (size_t)n)
                exit_with_msg("Error writing checkpoint rows.", -1);
            fclose(f);
        }
        printf("\t>Checkpoint rows written: %lld (every %lld rows) to %s\n\n", (long long)nrows, (long long)K,
               dump_dir);
    }
    if (json) {
        printf("{\"score\": %d, \"end\": [%lld, %lld], \"start\": [%lld, %lld], \"swapped\": %d, "
               "\"m\": %lld, \"n\": %lld, \"checkpoint_rows\": %lld, \"checkpoint_cols\": %lld, "
               "\"ms_forward\": %.3f, \"ms_end_search\": %.3f, \"ms_traceback\": %.3f, \"cigar\": \"%s\"}\n",
               res.score, (long long)res.end_row, (long long)res.end_col, (long long)res.start_row,
               (long long)res.start_col, res.swapped, (long long)res.m, (long long)res.n,
               (long long)res.checkpoint_rows, (long long)res.checkpoint_cols, res.ms_forward, res.ms_end_search,
               res.ms_traceback, res.cigar ? res.cigar : "");
    }
    swh_result_free(&res);
    swh_ctx_destroy(ctx);
    return 0;
}
