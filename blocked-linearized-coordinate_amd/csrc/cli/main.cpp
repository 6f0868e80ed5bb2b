// cpd64 -- the reference's command line (src/main.cpp:70-451) on the MI355X
// library.  Flags, defaults and the lines scripts parse are kept:
//   "--> Iter %3d: f = ... f-delta = ... time = ..., (s)", "Final fit: %0.9f",
//   "Total time for MTTKRP: %0.3f (time per iter: %0.3f ms)",
//   "Results of <name> is correct" / "... is incorrect by <count>".
// Additions: --precision f64|f32|bf16, --mode-order natural|longest, --graph,
// --device-budget BYTES;
// --max_block_size is accepted beside --max-block-size.  Without -k the
// reference runs its CPU ALTO path; this build runs the GPU path with the
// AUTO kernel (-k 10) and says so.
#include <getopt.h>
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "alto_dev.hpp"
#include "blco.hpp"
#include "blco_hip.h"
#include "blco_host.h"
#include "cpd_gpu.hpp"
#include "kruskal_model.hpp"
#include "sptensor.hpp"

static const struct option long_opt[] = {
    {"help", 0, NULL, 'h'},          {"version", 0, NULL, 'v'},
    {"input", 1, NULL, 'i'},         {"output", 1, NULL, 'o'},
    {"bin", 1, NULL, 'b'},           {"rank", 1, NULL, 'r'},
    {"max-iter", 1, NULL, 'm'},      {"seed", 1, NULL, 'x'},
    {"dims", 1, NULL, 'd'},          {"target-mode", 1, NULL, 't'},
    {"sparsity", 1, NULL, 's'},      {"epsilon", 1, NULL, 'e'},
    {"file", 1, NULL, 'f'},          {"check", 0, NULL, 'c'},
    {"bench", 0, NULL, 'p'},         {"kernel-id", 1, NULL, 'k'},
    {"num-partitions", 1, NULL, 'n'}, {"device", 1, NULL, 1001},
    {"thread-cf", 1, NULL, 1002},    {"stream-data", 0, NULL, 1003},
    {"max-block-size", 1, NULL, 1004}, {"max_block_size", 1, NULL, 1004},
    {"batch", 0, NULL, 1005},        {"precision", 1, NULL, 1006},
    {"mode-order", 1, NULL, 1007},   {"graph", 0, NULL, 1008},
    {"device-budget", 1, NULL, 1009},
    {NULL, 0, NULL, 0}};
static const char* short_opt = "hvi:o:b:r:m:x:d:t:s:e:f:cpk:n:";

static void Usage(const char* call)
{
    fprintf(stderr, "Usage: %s [OPTIONS]\n", call);
    fprintf(stderr,
            "Options:\n"
            "\t-h or --help         Display this information\n"
            "\t-v or --version      Display version information\n"
            "\t-i or --input        Input tensor file in text\n"
            "\t-o or --output       Output model to file\n"
            "\t-b or --bin          Input tensor file in binary\n"
            "\t-f or --file         Save tensor to another format\n"
            "\t-r or --rank         Rank\n"
            "\t-m or --max-iter     Maximum outer iterations\n"
            "\t-t or --target-mode  Target mode of tensor (-1: all modes)\n"
            "\t-e or --epsilon      Convergence criteria\n"
            "\t-x or --seed         Random value seed\n"
            "\t-d or --dims         Dimension lengths (I,J,K)\n"
            "\t-s or --sparsity     Sparsity of generated tensor\n"
            "\t-c or --check        Check GPU MTTKRP against a sequential COO MTTKRP\n"
            "\t-p or --bench        Run the GPU MTTKRP benchmark\n"
            "\t-k or --kernel-id    GPU kernel: 1 lvl1, 2 lvl2 (row window), 3 lvl3, 10 auto,\n"
            "\t                     11/12 batched, 13 atomic, 14 row-ordered, 20 selector\n"
            "\t-n or --num-partitions Number of lvl3 partial matrices\n"
            "\t--device             GPU ordinal\n"
            "\t--thread-cf          Thread coarsening factor (tiles of 64 per wave)\n"
            "\t--stream-data        Stream blocks to the GPU during computation\n"
            "\t--max-block-size     Block size (nonzeros) when streaming\n"
            "\t--batch              Launch all blocks at once\n"
            "\t--precision          f64 (default), f32 or bf16 factors\n"
            "\t--mode-order         natural (default) or longest\n"
            "\t--graph              Replay ALS sweeps from a captured hipGraph\n"
            "\t--device-budget      Device bytes for the tensor (suffix K, M or G: KiB, MiB,\n"
            "\t                     GiB); what does not fit streams in blocks of\n"
            "\t                     --max-block-size nonzeros\n");
}

// "32G", "512M", "4096": bytes, with binary K / M / G suffixes; 0 on error.
static uint64_t ParseBytes(const char* arg)
{
    char* end = nullptr;
    const double v = strtod(arg, &end);
    if (!end || end == arg || v <= 0.0) return 0;
    double mul = 1.0;
    if (*end == 'K' || *end == 'k') mul = 1024.0, ++end;
    else if (*end == 'M' || *end == 'm') mul = 1048576.0, ++end;
    else if (*end == 'G' || *end == 'g') mul = 1073741824.0, ++end;
    if (*end == 'B' || *end == 'b') ++end;
    if (*end) return 0;
    return (uint64_t)(v * mul);
}

static std::vector<IType> ParseDimensions(const char* arg)
{
    std::vector<IType> dims;
    std::string s(arg);
    size_t pos = 0;
    while (pos <= s.size()) {
        size_t end = s.find(',', pos);
        std::string tok = s.substr(pos, end == std::string::npos ? std::string::npos : end - pos);
        if (!tok.empty()) dims.push_back((IType)std::stoull(tok));
        if (end == std::string::npos) break;
        pos = end + 1;
    }
    return dims;
}

static void PrintTensorInfo(IType rank, int max_iters, SparseTensor* X)
{
    double cells = 1.0;
    for (int i = 0; i < X->nmodes; ++i) cells *= (double)X->dims[i];
    fprintf(stderr, "# Modes         = %u\n", X->nmodes);
    fprintf(stderr, "Rank            = %llu\n", rank);
    fprintf(stderr, "Sparsity        = %f\n", (double)X->nnz / cells);
    fprintf(stderr, "Max iters       = %d\n", max_iters);
    fprintf(stderr, "Dimensions      = [%llu", X->dims[0]);
    for (int i = 1; i < X->nmodes; ++i) fprintf(stderr, " X %llu", X->dims[i]);
    fprintf(stderr, "]\nNNZ             = %llu\n", X->nnz);
}

// Sequential COO MTTKRP, U[mode] += MTTKRP (the -c ground truth).
static void mttkrp_coo(SparseTensor* X, KruskalModel* M, int mode)
{
    const IType R = M->rank;
    std::vector<FType> row(R);
    for (IType i = 0; i < X->nnz; ++i) {
        for (IType r = 0; r < R; ++r) row[r] = X->vals[i];
        for (int m = 0; m < X->nmodes; ++m) {
            if (m == mode) continue;
            const FType* f = M->U[m] + X->cidx[m][i] * R;
            for (IType r = 0; r < R; ++r) row[r] *= f[r];
        }
        FType* o = M->U[mode] + X->cidx[mode][i] * R;
        for (IType r = 0; r < R; ++r) o[r] += row[r];
    }
}

static void VerifyResult(const char* name, const FType* truth, const FType* factor, IType size)
{
    IType cnt = 0;
    for (IType i = 0; i < size; ++i) {
        FType moe = std::max(std::fabs(truth[i]) * 1e-6, 1e-6);
        if (std::fabs(truth[i] - factor[i]) > moe) ++cnt;
    }
    if (cnt == 0) fprintf(stderr, "Results of %s is correct\n", name);
    else fprintf(stderr, "Results of %s is incorrect by %llu\n", name, cnt);
}

int main(int argc, char** argv)
{
    int max_iters = 10;
    IType rank = 16;
    std::vector<IType> dims;
    int target_mode = 0;
    std::string text_file, text_file_out, binary_file;
    double sparsity = 0.1, epsilon = 1e-5;
    int seed = (int)time(NULL);
    int save_to_file = 0;
    bool do_check = false, do_bench = false, stream_data = false, do_batching = false;
    int kernel_id = 0, num_partitions = 0, device = 0, precision = 0, mode_order = 0, graph = 0;
    IType thread_cf = 2, max_block_size = 16777216;
    uint64_t device_budget = 0;

    int c;
    while ((c = getopt_long(argc, argv, short_opt, long_opt, NULL)) != -1) {
        switch (c) {
            case 'h': Usage(argv[0]); return 0;
            case 'v': fprintf(stdout, "%s version 0.1.1 (gfx950)\n", argv[0]); return 0;
            case 'i': text_file = optarg; break;
            case 'o': text_file_out = optarg; break;
            case 'b': binary_file = optarg; break;
            case 'r': rank = (IType)atoll(optarg); if ((long long)rank < 1) { fprintf(stderr, "Invalid -rank: %s.\n", optarg); return -1; } break;
            case 'm': max_iters = atoi(optarg); if (max_iters < 0) { fprintf(stderr, "Invalid -max-iter: %s.\n", optarg); return -1; } break;
            case 'x': seed = atoi(optarg); break;
            case 'd': dims = ParseDimensions(optarg); if (dims.empty()) { fprintf(stderr, "Invalid -dims: %s.\n", optarg); return -1; } break;
            case 't': target_mode = atoi(optarg); if (target_mode < -1) { fprintf(stderr, "Invalid -target-mode: %s.\n", optarg); return -1; } break;
            case 's': sparsity = atof(optarg); if (sparsity <= 0.0) { fprintf(stderr, "Invalid -sparsity: %s.\n", optarg); return -1; } break;
            case 'e': epsilon = atof(optarg); if (epsilon < 0.0) { fprintf(stderr, "Invalid -epsilon: %s.\n", optarg); return -1; } break;
            case 'f': save_to_file = atoi(optarg); break;
            case 'c': do_check = true; break;
            case 'p': do_bench = true; break;
            case 'k': kernel_id = atoi(optarg); if (kernel_id < 0) { fprintf(stderr, "Invalid -kernel-id: %s.\n", optarg); return -1; } break;
            case 'n': num_partitions = atoi(optarg); if (num_partitions <= 0) { fprintf(stderr, "Invalid -num-partitions: %s.\n", optarg); return -1; } break;
            case 1001: device = atoi(optarg); break;
            case 1002: thread_cf = (IType)atoi(optarg); if ((int)thread_cf <= 0) { fprintf(stderr, "Invalid -thread-cf: %s.\n", optarg); return -1; } break;
            case 1003: stream_data = true; break;
            case 1004: max_block_size = (IType)atoll(optarg); if ((long long)max_block_size <= 0) { fprintf(stderr, "Invalid -max-block-size: %s.\n", optarg); return -1; } break;
            case 1005: do_batching = true; break;
            case 1006:
                if (!strcmp(optarg, "f64")) precision = 0;
                else if (!strcmp(optarg, "f32")) precision = 1;
                else if (!strcmp(optarg, "bf16")) precision = 2;
                else { fprintf(stderr, "Invalid --precision: %s.\n", optarg); return -1; }
                break;
            case 1007: mode_order = !strcmp(optarg, "longest") ? 1 : 0; break;
            case 1008: graph = 1; break;
            case 1009: device_budget = ParseBytes(optarg); if (!device_budget) { fprintf(stderr, "Invalid --device-budget: %s.\n", optarg); return -1; } break;
            default: Usage(argv[0]); return -1;
        }
    }
    const int known[] = {1, 2, 3, 10, 11, 12, 13, 14, 20};
    if (kernel_id && std::find(std::begin(known), std::end(known), kernel_id) == std::end(known)) {
        fprintf(stderr, "!ERROR! Unknown kernel %d\n", kernel_id);
        return -1;
    }

    SparseTensor* X = nullptr;
    if (!binary_file.empty()) {
        ImportSparseTensor(binary_file.c_str(), BINARY_FORMAT, &X);
        if (save_to_file) ExportSparseTensor(NULL, TEXT_FORMAT, X);
    } else if (!text_file.empty()) {
        ImportSparseTensor(text_file.c_str(), TEXT_FORMAT, &X);
        if (save_to_file) ExportSparseTensor(NULL, BINARY_FORMAT, X);
    } else if (dims.empty()) {
        fprintf(stderr, "No dims specified... exiting\n");
        Usage(argv[0]);
        return -1;
    } else {
        if (dims.size() > MAX_NUM_MODES) { fprintf(stderr, "at most %d modes\n", MAX_NUM_MODES); return -1; }
        blco_gen_spec sp;
        std::memset(&sp, 0, sizeof sp);
        sp.nmodes = (int)dims.size();
        double cells = 1.0;
        for (size_t n = 0; n < dims.size(); ++n) { sp.dims[n] = dims[n]; cells *= (double)dims[n]; }
        sp.target_nnz = (uint64_t)std::max(1.0, sparsity * cells);
        sp.gen_rank = (int)rank;
        sp.value_kind = 0;
        sp.seed = (uint64_t)seed;
        GenerateSparseTensor(&sp, &X);
        if (save_to_file) { ExportSparseTensor(NULL, TEXT_FORMAT, X); ExportSparseTensor(NULL, BINARY_FORMAT, X); }
    }
    if (target_mode >= X->nmodes) { fprintf(stderr, "Invalid -target-mode: %d.\n", target_mode); return -1; }

    if (kernel_id == 0) {
        fprintf(stderr, "note: the CPU ALTO path is not part of this build; running the GPU path (-k 10)\n");
        kernel_id = 10;
    }
    if (blco_device_count() <= device) {
        fprintf(stderr, "no GPU %d visible\n", device);
        return -1;
    }
    if (hipSetDevice(device) != hipSuccess) { fprintf(stderr, "hipSetDevice failed\n"); return -1; }
    if (do_check) max_iters = 1;
    cpals_set_device_budget(device_budget, max_block_size);
    if (!stream_data) max_block_size = X->nnz;
    if (num_partitions == 0) num_partitions = omp_get_max_threads();  // src/main.cpp:125
    cpals_set_precision(precision);
    cpals_set_mode_order(mode_order);
    cpals_set_graph(graph);
    blcotensor* tensor = gen_blcotensor_host<IType>(X, max_block_size);
    PrintTensorInfo(rank, max_iters, X);
    KruskalModel* M;
    CreateKruskalModel(X->nmodes, X->dims, rank, &M);
    KruskalModelRandomInit(M, (unsigned int)seed);

    if (do_check) {
        if (!do_bench) printf("Warning: check not implemented for CPD. Checking MTTKRP instead\n");
        std::vector<std::vector<FType>> truth(X->nmodes), init(X->nmodes);
        for (int m = 0; m < X->nmodes; ++m) init[m].assign(M->U[m], M->U[m] + X->dims[m] * rank);
        printf("===Do base run===\n");
        const int m0 = target_mode < 0 ? 0 : target_mode, m1 = target_mode < 0 ? X->nmodes - 1 : target_mode;
        for (int m = m0; m <= m1; ++m) mttkrp_coo(X, M, m);
        printf("--> MTTKRP sequential base run done.\n");
        for (int m = 0; m < X->nmodes; ++m) {
            truth[m].assign(M->U[m], M->U[m] + X->dims[m] * rank);
            std::memcpy(M->U[m], init[m].data(), sizeof(FType) * init[m].size());
        }
        printf("===Do ALTO GPU run===\n");
        mttkrp_alto_dev<IType>(tensor, M, kernel_id, max_iters, target_mode, thread_cf, stream_data,
                               do_batching, num_partitions);
        printf("MTTKRP ALTO GPU run done.\n");
        printf("===Verify ALTO-DEV MTTKRP=== (target mode: %d)\n", target_mode);
        for (int m = 0; m < X->nmodes; ++m) {
            printf("mode %d: ", m);
            fflush(stdout);
            VerifyResult("mttkrp_alto_dev", truth[m].data(), M->U[m], X->dims[m] * rank);
        }
    } else if (do_bench) {
        mttkrp_alto_dev<IType>(tensor, M, kernel_id, max_iters, target_mode, thread_cf, stream_data,
                               do_batching, num_partitions);
    } else {
        cpals_blco_dev(tensor, M, max_iters, epsilon, kernel_id, stream_data, do_batching, (int)thread_cf,
                       num_partitions);
        if (!text_file_out.empty()) {
            ExportKruskalModel(M, text_file_out.c_str());
            printf("--> Kruskal model saved to %s\n", text_file_out.c_str());
        }
    }
    DestroyKruskalModel(M);
    delete_blcotensor_host(tensor);
    DestroySparseTensor(X);
    return 0;
}
