// sptensor.cpp -- COO tensors and the .tns / .bin file formats
// (src/sptensor.cpp:22-297 of the reference define the layouts).
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <new>

#include "internal.hpp"

using namespace blco;

extern "C" {

int blco_coo_create(int nmodes, const uint64_t* dims, uint64_t nnz, const uint64_t* cidx,
                    const double* vals, blco_coo** out)
{
    if (!out || !dims || (nnz && (!cidx || !vals))) return fail(BLCO_EINVAL, "null argument");
    *out = nullptr;
    if (nmodes < 1 || nmodes > BLCO_MAX_MODES) return fail(BLCO_EINVAL, "nmodes %d", nmodes);
    blco_coo* t = new (std::nothrow) blco_coo();
    if (!t) return fail(BLCO_ENOMEM, "out of host memory");
    t->nmodes = nmodes;
    t->nnz = nnz;
    for (int n = 0; n < nmodes; ++n) {
        if (dims[n] == 0) { delete t; return fail(BLCO_EINVAL, "mode %d has length 0", n); }
        t->dims[n] = dims[n];
    }
    try {
        t->cidx.assign(cidx, cidx + (uint64_t)nmodes * nnz);
        t->vals.assign(vals, vals + nnz);
    } catch (const std::bad_alloc&) {
        delete t;
        return fail(BLCO_ENOMEM, "out of host memory");
    }
    for (int n = 0; n < nmodes; ++n)
        for (uint64_t i = 0; i < nnz; ++i)
            if (t->cidx[(uint64_t)n * nnz + i] >= dims[n]) {
                delete t;
                return fail(BLCO_EINVAL, "coordinate %llu of mode %d out of range",
                            (unsigned long long)i, n);
            }
    *out = t;
    return BLCO_OK;
}

static int import_bin(FILE* fp, blco_coo* t)
{
    uint64_t nm = 0;
    if (fread(&nm, 8, 1, fp) != 1 || nm < 1 || nm > BLCO_MAX_MODES)
        return fail(BLCO_EINVAL, "bad .bin header (nmodes)");
    t->nmodes = (int)nm;
    if (fread(t->dims, 8, nm, fp) != nm || fread(&t->nnz, 8, 1, fp) != 1)
        return fail(BLCO_EINVAL, "truncated .bin header");
    try {
        t->cidx.resize(nm * t->nnz);
        t->vals.resize(t->nnz);
    } catch (const std::bad_alloc&) {
        return fail(BLCO_ENOMEM, "out of host memory");
    }
    if (fread(t->cidx.data(), 8, nm * t->nnz, fp) != nm * t->nnz ||
        fread(t->vals.data(), 8, t->nnz, fp) != t->nnz)
        return fail(BLCO_EINVAL, "truncated .bin body");
    for (uint64_t n = 0; n < nm; ++n)
        for (uint64_t i = 0; i < t->nnz; ++i)
            if (t->cidx[n * t->nnz + i] >= t->dims[n])
                return fail(BLCO_EINVAL, ".bin coordinate out of range");
    return BLCO_OK;
}

// .tns: one nonzero per line, 1-based indices then the value; blank lines and
// lines starting with '#' are skipped; mode lengths are the largest index
// seen per mode (read_tns_dims, src/sptensor.cpp:74-131).
static int import_tns(FILE* fp, blco_coo* t)
{
    char* line = nullptr;
    size_t cap = 0;
    ssize_t len;
    int nm = 0;
    std::vector<uint64_t> idx;  // nonzero-major while reading
    while ((len = getline(&line, &cap, fp)) != -1) {
        if (len <= 1 || line[0] == '#') continue;
        char* p = line;
        uint64_t row[BLCO_MAX_MODES + 1];
        double v = 0.0;
        int cnt = 0;
        // count tokens on the first data line
        if (nm == 0) {
            char* q = line;
            int tok = 0;
            while (*q) {
                while (*q == ' ' || *q == '\t' || *q == '\n' || *q == '\r') ++q;
                if (!*q) break;
                ++tok;
                while (*q && *q != ' ' && *q != '\t' && *q != '\n' && *q != '\r') ++q;
            }
            nm = tok - 1;
            if (nm < 1 || nm > BLCO_MAX_MODES) { free(line); return fail(BLCO_EINVAL, ".tns has %d modes", nm); }
            t->nmodes = nm;
        }
        for (cnt = 0; cnt < nm; ++cnt) {
            char* e;
            errno = 0;
            unsigned long long x = strtoull(p, &e, 10);
            if (e == p || x == 0) { free(line); return fail(BLCO_EINVAL, "bad .tns index"); }
            row[cnt] = x - 1;
            p = e;
        }
        char* e;
        v = strtod(p, &e);
        if (e == p) { free(line); return fail(BLCO_EINVAL, "bad .tns value"); }
        for (int n = 0; n < nm; ++n) {
            idx.push_back(row[n]);
            if (row[n] + 1 > t->dims[n]) t->dims[n] = row[n] + 1;
        }
        t->vals.push_back(v);
    }
    free(line);
    if (nm == 0) return fail(BLCO_EINVAL, "empty .tns file");
    t->nnz = t->vals.size();
    t->cidx.resize((uint64_t)nm * t->nnz);
    for (uint64_t i = 0; i < t->nnz; ++i)
        for (int n = 0; n < nm; ++n) t->cidx[(uint64_t)n * t->nnz + i] = idx[i * nm + n];
    return BLCO_OK;
}

int blco_coo_import(const char* path, int format, blco_coo** out)
{
    if (!path || !out) return fail(BLCO_EINVAL, "null argument");
    *out = nullptr;
    FILE* fp = fopen(path, format == 1 ? "rb" : "r");
    if (!fp) return fail(BLCO_EINVAL, "cannot open %s", path);
    blco_coo* t = new (std::nothrow) blco_coo();
    int rc = t ? (format == 1 ? import_bin(fp, t) : import_tns(fp, t)) : BLCO_ENOMEM;
    fclose(fp);
    if (rc) { delete t; return rc; }
    *out = t;
    return BLCO_OK;
}

int blco_coo_export(const blco_coo* t, const char* path, int format)
{
    if (!t || !path) return fail(BLCO_EINVAL, "null argument");
    FILE* fp = fopen(path, format == 1 ? "wb" : "w");
    if (!fp) return fail(BLCO_EINVAL, "cannot open %s", path);
    if (format == 1) {
        uint64_t nm = (uint64_t)t->nmodes;
        fwrite(&nm, 8, 1, fp);
        fwrite(t->dims, 8, nm, fp);
        fwrite(&t->nnz, 8, 1, fp);
        fwrite(t->cidx.data(), 8, nm * t->nnz, fp);
        fwrite(t->vals.data(), 8, t->nnz, fp);
    } else {
        for (uint64_t i = 0; i < t->nnz; ++i) {
            for (int n = 0; n < t->nmodes; ++n)
                fprintf(fp, "%llu ", (unsigned long long)(t->cidx[(uint64_t)n * t->nnz + i] + 1));
            fprintf(fp, "%g\n", t->vals[i]);
        }
    }
    int bad = ferror(fp);
    fclose(fp);
    return bad ? fail(BLCO_EINVAL, "write to %s failed", path) : BLCO_OK;
}

int blco_coo_info(const blco_coo* t, int* nmodes, uint64_t* dims, uint64_t* nnz)
{
    if (!t) return fail(BLCO_EINVAL, "null tensor");
    if (nmodes) *nmodes = t->nmodes;
    if (dims) std::memcpy(dims, t->dims, sizeof(uint64_t) * t->nmodes);
    if (nnz) *nnz = t->nnz;
    return BLCO_OK;
}

const uint64_t* blco_coo_cidx(const blco_coo* t) { return t ? t->cidx.data() : nullptr; }
const double* blco_coo_vals(const blco_coo* t) { return t ? t->vals.data() : nullptr; }
void blco_coo_free(blco_coo* t) { delete t; }

}  // extern "C"
