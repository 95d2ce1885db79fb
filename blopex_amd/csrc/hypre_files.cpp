// Readers of the HYPRE IJ matrix and ParVector file sets (formats described
// in include/hypre_files.h): parsed on the host, then handed to the device
// constructors.  Numbers are parsed from a whole-file buffer with strtol /
// strtod, so a file set of ten million entries reads in seconds.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"
#include "hypre_files.h"

using namespace bx;

namespace {

bool slurp(const std::string &path, std::string &out) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    out.clear();
    char buf[1 << 16];
    size_t got;
    while ((got = fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, got);
    fclose(f);
    return true;
}

bool exists(const std::string &path) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    fclose(f);
    return true;
}

// whitespace-separated numbers of a buffer, one at a time
struct Scanner {
    const char *p, *end;
    const std::string *name;
    long line = 1;
    void skip() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == ',')) {
            if (*p == '\n') line++;
            p++;
        }
    }
    bool done() {
        skip();
        return p >= end;
    }
    bool integer(long *v) {
        skip();
        char *e;
        errno = 0;
        const long x = strtol(p, &e, 10);
        if (e == p || errno) return false;
        p = e;
        *v = x;
        return true;
    }
    bool real(double *v) {
        skip();
        char *e;
        const double x = strtod(p, &e);
        if (e == p) return false;
        p = e;
        *v = x;
        return true;
    }
};

std::string part_name(const char *prefix, int p) {
    char s[16];
    snprintf(s, sizeof s, ".%05d", p);
    return std::string(prefix) + s;
}

}  // namespace

extern "C" {

void hip_IJFree(void *p) { free(p); }

int hip_IJReadCSR(const char *prefix, long *n_out, long *nnz_out, int **rowptr_out,
                  int **colind_out, double **val_out) {
    if (!prefix || !n_out || !nnz_out || !rowptr_out || !colind_out || !val_out) {
        set_error("hip_IJReadCSR: null argument");
        return -1;
    }
    struct Entry {
        long i, j;
        double v;
    };
    std::vector<Entry> ent;
    long n = 0;
    int parts = 0;
    std::string text;
    for (;; parts++) {
        const std::string name = part_name(prefix, parts);
        if (!exists(name)) break;
        if (!slurp(name, text)) {
            set_error("hip_IJReadCSR: cannot read %s", name.c_str());
            return -1;
        }
        Scanner sc{text.data(), text.data() + text.size(), &name};
        long il, iu, jl, ju;
        if (!sc.integer(&il) || !sc.integer(&iu) || !sc.integer(&jl) || !sc.integer(&ju) ||
            il < 0 || iu < il - 1) {
            set_error("hip_IJReadCSR: %s: bad header line", name.c_str());
            return -1;
        }
        n = std::max(n, iu + 1);
        while (!sc.done()) {
            long i, j;
            double v;
            if (!sc.integer(&i) || !sc.integer(&j) || !sc.real(&v)) {
                set_error("hip_IJReadCSR: %s line %ld: expected \"i j value\"", name.c_str(),
                          sc.line);
                return -1;
            }
            if (i < il || i > iu) {
                set_error("hip_IJReadCSR: %s line %ld: row %ld outside the part's rows %ld..%ld",
                          name.c_str(), sc.line, i, il, iu);
                return -1;
            }
            ent.push_back({i, j, v});
        }
    }
    if (parts == 0) {
        set_error("hip_IJReadCSR: no file %s", part_name(prefix, 0).c_str());
        return -1;
    }
    if (n > 0x7fffffffL) {
        set_error("hip_IJReadCSR: %ld rows exceed int32 indices", n);
        return -1;
    }
    for (const Entry &e : ent)
        if (e.j < 0 || e.j >= n) {
            set_error("hip_IJReadCSR: column %ld outside the %ld x %ld matrix", e.j, n, n);
            return -1;
        }
    std::stable_sort(ent.begin(), ent.end(), [](const Entry &a, const Entry &b) {
        return a.i != b.i ? a.i < b.i : a.j < b.j;
    });
    // repeated (i, j): summed in file order
    size_t m = 0;
    for (size_t t = 0; t < ent.size(); t++) {
        if (m > 0 && ent[m - 1].i == ent[t].i && ent[m - 1].j == ent[t].j)
            ent[m - 1].v += ent[t].v;
        else
            ent[m++] = ent[t];
    }
    ent.resize(m);
    if ((long)m > 0x7fffffffL) {
        set_error("hip_IJReadCSR: %zu entries exceed int32 row pointers", m);
        return -1;
    }
    int *rowptr = (int *)malloc(sizeof(int) * (n + 1));
    int *colind = (int *)malloc(sizeof(int) * (m > 0 ? m : 1));
    double *val = (double *)malloc(sizeof(double) * (m > 0 ? m : 1));
    if (!rowptr || !colind || !val) {
        free(rowptr);
        free(colind);
        free(val);
        set_error("hip_IJReadCSR: out of host memory");
        return -1;
    }
    for (long r = 0; r <= n; r++) rowptr[r] = 0;
    for (size_t t = 0; t < m; t++) {
        rowptr[ent[t].i + 1]++;
        colind[t] = (int)ent[t].j;
        val[t] = ent[t].v;
    }
    for (long r = 0; r < n; r++) rowptr[r + 1] += rowptr[r];
    *n_out = n;
    *nnz_out = (long)m;
    *rowptr_out = rowptr;
    *colind_out = colind;
    *val_out = val;
    return 0;
}

int hip_ParVectorsRead(const char *prefix, long *n_out, int *m_out, double **v_out) {
    if (!prefix || !n_out || !m_out || !v_out) {
        set_error("hip_ParVectorsRead: null argument");
        return -1;
    }
    auto name = [&](int j, int i) {
        return std::string(prefix) + "." + std::to_string(j) + "." + std::to_string(i);
    };
    int m = 0, parts = 0;
    while (exists(name(m, 0))) m++;
    while (exists(name(0, parts))) parts++;
    if (m == 0) {
        set_error("hip_ParVectorsRead: no file %s", name(0, 0).c_str());
        return -1;
    }
    std::vector<std::vector<double>> cols(m);
    std::string text;
    for (int j = 0; j < m; j++) {
        for (int i = 0; i < parts; i++) {
            const std::string nm = name(j, i);
            if (!slurp(nm, text)) {
                set_error("hip_ParVectorsRead: missing file %s", nm.c_str());
                return -1;
            }
            Scanner sc{text.data(), text.data() + text.size(), &nm};
            long rows;
            if (!sc.integer(&rows) || rows < 0) {
                set_error("hip_ParVectorsRead: %s: bad row count", nm.c_str());
                return -1;
            }
            for (long r = 0; r < rows; r++) {
                double v;
                if (!sc.real(&v)) {
                    set_error("hip_ParVectorsRead: %s: %ld of %ld values", nm.c_str(), r, rows);
                    return -1;
                }
                cols[j].push_back(v);
            }
            if (!sc.done()) {
                set_error("hip_ParVectorsRead: %s: more than %ld values", nm.c_str(), rows);
                return -1;
            }
        }
        if (cols[j].size() != cols[0].size()) {
            set_error("hip_ParVectorsRead: vector %d has %zu rows, vector 0 %zu", j,
                      cols[j].size(), cols[0].size());
            return -1;
        }
    }
    const long n = (long)cols[0].size();
    double *v = (double *)malloc(sizeof(double) * (n * m > 0 ? n * m : 1));
    if (!v) {
        set_error("hip_ParVectorsRead: out of host memory");
        return -1;
    }
    for (int j = 0; j < m; j++) std::copy(cols[j].begin(), cols[j].end(), v + (size_t)j * n);
    *n_out = n;
    *m_out = m;
    *v_out = v;
    return 0;
}

hip_CSRMatrix *hip_CSRMatrixReadIJ(const char *prefix) {
    long n, nnz;
    int *rowptr, *colind;
    double *val;
    if (hip_IJReadCSR(prefix, &n, &nnz, &rowptr, &colind, &val) != 0) return nullptr;
    hip_CSRMatrix *A = hip_CSRMatrixCreate(n, nnz, rowptr, colind, val);
    free(rowptr);
    free(colind);
    free(val);
    return A;
}

hip_Multi_Vector *hip_Multi_VectorReadParVectors(const char *prefix) {
    long n;
    int m;
    double *v;
    if (hip_ParVectorsRead(prefix, &n, &m, &v) != 0) return nullptr;
    hip_Multi_Vector *x = hip_Multi_VectorCreate(n, m, HIPMV_FP64);
    hip_Multi_VectorUpload(x, v);
    free(v);
    return x;
}

}  // extern "C"
