// Catalog readers (host code): fast-food and ascii/csv column files, behind the
// C entry points of include/corrfunc_io.h. Follows the reference's
// io/io.c:28-284 (formats, checks, skip rules) and io/ftread.c:34-67 (record
// framing); written for this library, without its fixed-size buffers.
#include <cerrno>
#include <cinttypes>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <spawn.h>
#include <sys/stat.h>
#include <sys/wait.h>
#include <unistd.h>

#include "../../include/corrfunc_io.h"

extern char** environ;

namespace {

bool file_exists(const std::string& path)
{
    struct stat st;
    return stat(path.c_str(), &st) == 0;
}

// The reference (io/io.c:60-97) reads `name` if it exists, else uncompresses
// `name.gz` in place with gunzip and reads the result.
bool locate(const char* filename)
{
    if (file_exists(filename)) return true;
    const std::string gz = std::string(filename) + ".gz";
    if (!file_exists(gz)) {
        fprintf(stderr, "ERROR: Could not find file: neither as `%s' nor as `%s'\n", filename, gz.c_str());
        return false;
    }
    fprintf(stderr, "Could not locate `%s' but found the gzip file `%s'. Running `gunzip %s' now to uncompress\n",
            filename, gz.c_str(), gz.c_str());
    char arg0[] = "gunzip";
    std::vector<char> arg1(gz.begin(), gz.end());
    arg1.push_back('\0');
    char* argv[] = {arg0, arg1.data(), nullptr};
    pid_t pid;
    if (posix_spawnp(&pid, "gunzip", nullptr, nullptr, argv, environ) != 0) {
        fprintf(stderr, "ERROR: could not start gunzip for `%s'\n", gz.c_str());
        return false;
    }
    int status = 0;
    while (waitpid(pid, &status, 0) < 0) {
        if (errno != EINTR) return false;
    }
    if (!WIFEXITED(status) || WEXITSTATUS(status) != 0 || !file_exists(filename)) {
        fprintf(stderr, "ERROR: gunzip of `%s' failed\n", gz.c_str());
        return false;
    }
    return true;
}

// One Fortran unformatted record: a 4-byte length, the payload, the length
// again; both lengths must equal `bytes` (io/ftread.c:34-67).
bool read_record(FILE* fp, void* dst, size_t bytes, const char* what)
{
    unsigned n1 = 0, n2 = 0;
    if (fread(&n1, sizeof n1, 1, fp) != 1) {
        fprintf(stderr, "ERROR: fast-food read of %s: file empty or truncated\n", what);
        return false;
    }
    if (bytes > 0 && fread(dst, 1, bytes, fp) != bytes) {
        fprintf(stderr, "ERROR: fast-food read of %s: %zu bytes expected, file too short\n", what, bytes);
        return false;
    }
    if (fread(&n2, sizeof n2, 1, fp) != 1) {
        fprintf(stderr, "ERROR: fast-food read of %s: file too short\n", what);
        return false;
    }
    if (n1 != n2 || (size_t)n1 != bytes) {
        fprintf(stderr, "ERROR: fast-food read of %s: record markers %u and %u, %zu bytes expected\n", what, n1, n2,
                bytes);
        return false;
    }
    return true;
}

struct Columns {
    std::vector<void*> p;
    explicit Columns(int n) : p(n, nullptr) {}
    ~Columns()
    {
        for (void* q : p) free(q);
    }
    bool alloc(size_t bytes)
    {
        for (auto& q : p) {
            q = malloc(bytes > 0 ? bytes : 1);
            if (!q) {
                fprintf(stderr, "ERROR: could not allocate %zu bytes for a catalog column\n", bytes);
                return false;
            }
        }
        return true;
    }
    void release(void** out)
    {
        for (size_t i = 0; i < p.size(); i++) {
            out[i] = p[i];
            p[i] = nullptr;
        }
    }
};

int64_t read_fastfood(const char* filename, size_t size, int num_fields, void** data)
{
    FILE* fp = fopen(filename, "rb");
    if (!fp) {
        fprintf(stderr, "ERROR: could not open `%s': %s\n", filename, strerror(errno));
        return -1;
    }
    struct Closer {
        FILE* f;
        ~Closer() { fclose(f); }
    } closer{fp};
    int idat[5];
    if (!read_record(fp, idat, sizeof idat, "the header (idat)")) return -1;
    const int64_t np = idat[1];
    if (np < 0) {
        fprintf(stderr, "ERROR: fast-food file `%s' declares %" PRId64 " particles\n", filename, np);
        return -1;
    }
    Columns cols(num_fields);
    if (!cols.alloc(size * (size_t)np)) return -1;
    if (np > 0) {
        // skip the fdat (9 floats) and znow (1 float) records
        if (fseek(fp, (long)(4 + 36 + 4 + 4 + 4 + 4), SEEK_CUR) != 0) {
            fprintf(stderr, "ERROR: fast-food file `%s' ends inside its header\n", filename);
            return -1;
        }
        unsigned marker = 0;
        if (fread(&marker, sizeof marker, 1, fp) != 1 || fseek(fp, -4, SEEK_CUR) != 0) {
            fprintf(stderr, "ERROR: fast-food file `%s' has no data records\n", filename);
            return -1;
        }
        const size_t fsize = marker / (uint64_t)np;
        if (fsize != 4 && fsize != 8) {
            fprintf(stderr, "ERROR: data-type in file = %zu must be either 4 byte (float) or 8 byte (double)\n",
                    fsize);
            return -1;
        }
        if (fsize != size)
            fprintf(stderr, "WARNING: File was written in a different precision than requested (file precision = "
                            "%zu requested precision = %zu)\n", fsize, size);
        std::vector<unsigned char> tmp(fsize == size ? 0 : fsize * (size_t)np);
        for (int k = 0; k < num_fields; k++) {
            char what[48];
            snprintf(what, sizeof what, "column %d", k);
            if (fsize == size) {
                if (!read_record(fp, cols.p[k], size * (size_t)np, what)) return -1;
            } else {
                if (!read_record(fp, tmp.data(), tmp.size(), what)) return -1;
                if (fsize == 4) {
                    const float* s = (const float*)tmp.data();
                    double* d = (double*)cols.p[k];
                    for (int64_t i = 0; i < np; i++) d[i] = (double)s[i];
                } else {
                    const double* s = (const double*)tmp.data();
                    float* d = (float*)cols.p[k];
                    for (int64_t i = 0; i < np; i++) d[i] = (float)s[i];
                }
            }
        }
    }
    cols.release(data);
    return np;
}

int64_t read_text(const char* filename, size_t size, int num_fields, void** data)
{
    FILE* fp = fopen(filename, "r");
    if (!fp) {
        fprintf(stderr, "ERROR: could not open `%s': %s\n", filename, strerror(errno));
        return -1;
    }
    std::vector<std::vector<double>> v(num_fields);
    std::vector<double> row(num_fields);
    char* line = nullptr;
    size_t cap = 0;
    int64_t n = 0;
    static const char delim[] = " ,\t";
    while (getline(&line, &cap, fp) >= 0) {
        char* save = nullptr;
        char* s = line;
        bool ok = true;
        for (int j = 0; j < num_fields; j++, s = nullptr) {
            char* tok = strtok_r(s, delim, &save);
            char* end = tok;
            if (tok) row[j] = strtod(tok, &end);
            if (!tok || end == tok) {
                ok = false;
                break;
            }
        }
        if (!ok) {
            fprintf(stderr, "io> WARNING: Could not parse all requested %d fields in line %" PRId64
                            ". Skipping that line\n", num_fields, n);
            continue;
        }
        for (int j = 0; j < num_fields; j++) v[j].push_back(row[j]);
        n++;
    }
    free(line);
    fclose(fp);
    Columns cols(num_fields);
    if (!cols.alloc(size * (size_t)n)) return -1;
    for (int j = 0; j < num_fields; j++) {
        if (size == 8) {
            memcpy(cols.p[j], v[j].data(), sizeof(double) * (size_t)n);
        } else {
            float* d = (float*)cols.p[j];
            for (int64_t i = 0; i < n; i++) d[i] = (float)v[j][i];
        }
    }
    cols.release(data);
    return n;
}

}  // namespace

extern "C" int64_t read_columns_into_array(const char* filename, const char* format, const size_t size,
                                           const int num_fields, void** data)
{
    if (num_fields < 1) {
        fprintf(stderr, "ERROR: Number of fields to read-in = %d must be at least 1\n", num_fields);
        return -1;
    }
    if (size != 4 && size != 8) {
        fprintf(stderr, "ERROR: Size of fields = %zu must be either 4 or 8\n", size);
        return -1;
    }
    if (!filename || !format || !data) {
        fprintf(stderr, "ERROR: read_columns_into_array needs a file name, a format and an output array\n");
        return -1;
    }
    if (format[0] != 'f' && format[0] != 'a' && format[0] != 'c') {
        fprintf(stderr, "ERROR: In %s> Unknown format `%s'\n", __func__, format);
        return -1;
    }
    if (!locate(filename)) return -1;
    return format[0] == 'f' ? read_fastfood(filename, size, num_fields, data)
                            : read_text(filename, size, num_fields, data);
}

extern "C" int64_t read_positions(const char* filename, const char* format, const size_t size, const int num_fields,
                                  ...)
{
    if (num_fields < 1) {
        fprintf(stderr, "ERROR: Number of fields to read-in = %d must be at least 1\n", num_fields);
        return -1;
    }
    std::vector<void*> cols(num_fields, nullptr);
    const int64_t np = read_columns_into_array(filename, format, size, num_fields, cols.data());
    if (np < 0) return np;
    va_list ap;
    va_start(ap, num_fields);
    for (int i = 0; i < num_fields; i++) {
        void** dst = va_arg(ap, void**);
        *dst = cols[i];
    }
    va_end(ap);
    return np;
}
