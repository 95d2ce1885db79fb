// hip_Multi_Vector: the device block behind the interpreter table, its
// interpreter slots (the multivector part of mv_InterfaceInterpreter) and the
// CSR / diagonal operator callbacks.
//
// Slot semantics follow the serial backend (multi_vector.c): every operation
// touches only the active columns, in ascending order; Gram results land in
// caller (host) matrices with leading dimension gh; MultiVecMat overwrites.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"
#include "fortran_matrix.h"
#include "kernels.h"

using namespace bx;

namespace {

hip_LobpcgSettings default_settings() {
    hip_LobpcgSettings s;
    s.mode = HIPLOBPCG_FP64;
    s.switch_factor = 8.0;
    s.stall_iterations = 50;
    s.refine_iterations = 0;
    s.profile = 0;
    s.refresh_interval = 20;
    s.final_rr = 0;
    s.guard_columns = 0;
    s.eig_rtol = 0.0;
    s.cond_restart = 0;
    s.size_guard = 0;
    s.guard_seed = 7;
    return s;
}

hip_Multi_Vector *alloc_header(long size, BlopexInt k, int precision) {
    BX_REQUIRE(size >= 0 && k >= 0, "hip_Multi_VectorCreate: bad shape %ld x %d", size, (int)k);
    BX_REQUIRE(precision == HIPMV_FP64 || precision == HIPMV_FP32 || precision == HIPMV_C64,
               "hip_Multi_VectorCreate: bad precision %d", precision);
    hip_Multi_Vector *x = (hip_Multi_Vector *)calloc(1, sizeof(hip_Multi_Vector));
    x->size = size;
    x->ld = size;
    x->num_vectors = k;
    x->precision = precision;
    x->active_indices = (BlopexInt *)malloc(sizeof(BlopexInt) * (k > 0 ? k : 1));
    for (BlopexInt i = 0; i < k; i++) x->active_indices[i] = i;
    x->num_active_vectors = k;
    x->settings = default_settings();
    return x;
}

inline hip_Multi_Vector *H(void *p) { return (hip_Multi_Vector *)p; }

void same_shape(const hip_Multi_Vector *x, const hip_Multi_Vector *y, const char *op) {
    BX_REQUIRE(x->size == y->size, "%s: row counts differ (%ld vs %ld)", op, x->size, y->size);
    BX_REQUIRE(x->precision == y->precision, "%s: precisions differ", op);
}

// The block as the FP64 kernels see it: a komplex column is 2 * size doubles
// (copies, clears, real axpy and the drand48 fill act on those directly).
struct View {
    int prec;
    long n, ld;
};
View view(const hip_Multi_Vector *x) {
    if (x->precision == HIPMV_C64) return {HIPMV_FP64, 2 * x->size, 2 * x->ld};
    return {x->precision, x->size, x->ld};
}
bool cplx(const hip_Multi_Vector *x) { return x->precision == HIPMV_C64; }
void real_only(const hip_Multi_Vector *x, const char *op) {
    BX_REQUIRE(x->precision != HIPMV_C64, "%s: complex blocks take HipDenseMatMultiVec", op);
}

// active positions of a caller mask over n entries (NULL: all)
std::vector<int> mask_positions(const BlopexInt *mask, BlopexInt n) {
    std::vector<int> p;
    for (BlopexInt i = 0; i < n; i++)
        if (!mask || mask[i]) p.push_back((int)i);
    return p;
}

}  // namespace

extern "C" {

hip_Multi_Vector *hip_Multi_VectorCreate(long size, BlopexInt k, int precision) {
    hip_Multi_Vector *x = alloc_header(size, k, precision);
    size_t bytes = (size_t)size * (size_t)k * elem_size(precision);
    if (bytes) {
        stream();
        BX_HIP(hipMalloc(&x->data, bytes));
    }
    x->owns_data = 1;
    return x;
}

hip_Multi_Vector *hip_Multi_VectorWrapDevice(void *data, long size, BlopexInt k, int precision) {
    hip_Multi_Vector *x = alloc_header(size, k, precision);
    x->data = data;
    x->owns_data = 0;
    return x;
}

void hip_Multi_VectorDestroy(hip_Multi_Vector *x) {
    if (!x) return;
    if (x->owns_data && x->data) {
        BX_HIP(hipStreamSynchronize(stream()));
        BX_HIP(hipFree(x->data));
    }
    free(x->active_indices);
    free(x);
}

int hip_Multi_VectorUpload(hip_Multi_Vector *x, const double *host) {
    const View v = view(x);
    size_t n = (size_t)v.n * x->num_vectors;
    if (!n) return 0;
    if (v.prec == HIPMV_FP64 && v.ld == v.n) {
        BX_HIP(hipMemcpyAsync(x->data, host, n * 8, hipMemcpyHostToDevice, stream()));
    } else {
        void *tmp = scratch(4, n * 8);
        BX_HIP(hipMemcpyAsync(tmp, host, n * 8, hipMemcpyHostToDevice, stream()));
        convert_block(v.n, x->num_vectors, tmp, v.n, HIPMV_FP64, x->data, v.ld, v.prec);
    }
    BX_HIP(hipStreamSynchronize(stream()));
    return 0;
}

int hip_Multi_VectorDownload(const hip_Multi_Vector *x, double *host) {
    const View v = view(x);
    size_t n = (size_t)v.n * x->num_vectors;
    if (!n) return 0;
    if (v.prec == HIPMV_FP64 && v.ld == v.n) {
        BX_HIP(hipMemcpyAsync(host, x->data, n * 8, hipMemcpyDeviceToHost, stream()));
    } else {
        void *tmp = scratch(4, n * 8);
        convert_block(v.n, x->num_vectors, x->data, v.ld, v.prec, tmp, v.n, HIPMV_FP64);
        BX_HIP(hipMemcpyAsync(host, tmp, n * 8, hipMemcpyDeviceToHost, stream()));
    }
    BX_HIP(hipStreamSynchronize(stream()));
    return 0;
}

int hip_Multi_VectorSetRandomValues(hip_Multi_Vector *x, BlopexInt seed) {
    // komplex: real and imaginary parts take consecutive values of the stream
    // (blopex_serial_complex/multivector/multi_vector.c:166-190)
    const View v = view(x);
    drand48_cols(v.prec, v.n, (int)seed, x->data, v.ld, colmap_of(x));
    return 0;
}

void hip_LobpcgSettingsDefault(hip_LobpcgSettings *s) {
    if (s) *s = default_settings();
}

int hip_Multi_VectorSetSettings(hip_Multi_Vector *x, const hip_LobpcgSettings *s) {
    if (!x || !s) return 1;
    if (s->mode != HIPLOBPCG_FP64 && s->mode != HIPLOBPCG_MIXED) {
        set_error("unknown solver mode %d", s->mode);
        return 1;
    }
    if (s->guard_columns < 0 || !(s->eig_rtol >= 0.0)) {
        set_error("guard_columns must be >= 0 and eig_rtol >= 0 (got %d, %g)", s->guard_columns,
                  s->eig_rtol);
        return 1;
    }
    x->settings = *s;
    return 0;
}

int hip_Multi_VectorGetStats(const hip_Multi_Vector *x, hip_LobpcgStats *s) {
    if (!x || !s) return 1;
    *s = x->stats;
    return 0;
}

// ---- interpreter slots --------------------------------------------------

static void *hipmv_CopyCreate(void *src_, BlopexInt copyValues) {
    hip_Multi_Vector *src = H(src_);
    hip_Multi_Vector *dst = hip_Multi_VectorCreate(src->size, src->num_vectors, src->precision);
    dst->settings = src->settings;
    const View vs = view(src), vd = view(dst);
    if (copyValues && src->size && src->num_vectors)
        convert_block(vs.n, src->num_vectors, src->data, vs.ld, vs.prec, dst->data, vd.ld, vd.prec);
    return dst;
}

static void hipmv_Destroy(void *x) { hip_Multi_VectorDestroy(H(x)); }

static BlopexInt hipmv_Width(void *x) { return H(x)->num_vectors; }

static BlopexInt hipmv_Height(void *x) { return (BlopexInt)H(x)->size; }

static void hipmv_SetMask(void *x_, BlopexInt *mask) {
    hip_Multi_Vector *x = H(x_);
    x->num_active_vectors = 0;
    for (BlopexInt i = 0; i < x->num_vectors; i++)
        if (!mask || mask[i]) x->active_indices[x->num_active_vectors++] = i;
}

static void hipmv_Copy(void *x_, void *y_) {
    hip_Multi_Vector *x = H(x_), *y = H(y_);
    same_shape(x, y, "CopyMultiVector");
    const View vx = view(x), vy = view(y);
    copy_cols(vx.prec, vx.n, x->data, vx.ld, colmap_of(x), y->data, vy.ld, colmap_of(y));
}

static void hipmv_Clear(void *x_) {
    hip_Multi_Vector *x = H(x_);
    const View v = view(x);
    clear_cols(v.prec, v.n, x->data, v.ld, colmap_of(x));
}

static void hipmv_SetRandom(void *x, BlopexInt seed) { hip_Multi_VectorSetRandomValues(H(x), seed); }

static void hipmv_InnerProd(void *x_, void *y_, BlopexInt gh, BlopexInt h, BlopexInt w, void *v) {
    hip_Multi_Vector *x = H(x_), *y = H(y_);
    same_shape(x, y, "MultiInnerProd");
    BX_REQUIRE(x->num_active_vectors == h && y->num_active_vectors == w,
               "MultiInnerProd: %d x %d result for %d x %d active columns", (int)h, (int)w,
               (int)x->num_active_vectors, (int)y->num_active_vectors);
    if (h == 0 || w == 0) return;
    if (cplx(x)) {
        komplex *g = (komplex *)scratch(3, (size_t)h * w * sizeof(komplex));
        zgram_cols(x->size, x->data, x->ld, colmap_of(x), y->data, y->ld, colmap_of(y), g, h);
        std::vector<komplex> tmp((size_t)h * w);
        download_sync(tmp.data(), g, tmp.size() * sizeof(komplex));
        for (BlopexInt j = 0; j < w; j++)
            memcpy((komplex *)v + (size_t)j * gh, tmp.data() + (size_t)j * h, sizeof(komplex) * h);
        return;
    }
    double *g = (double *)scratch(3, (size_t)h * w * sizeof(double));
    gram_cols(x->precision, x->size, x->data, x->ld, colmap_of(x), y->data, y->ld,
              colmap_of(y), g, h);
    std::vector<double> tmp((size_t)h * w);
    download_sync(tmp.data(), g, tmp.size() * sizeof(double));
    double *out = (double *)v;
    for (BlopexInt j = 0; j < w; j++)
        memcpy(out + (size_t)j * gh, tmp.data() + (size_t)j * h, sizeof(double) * h);
}

static void hipmv_InnerProdDiag(void *x_, void *y_, BlopexInt *mask, BlopexInt n, void *diag) {
    hip_Multi_Vector *x = H(x_), *y = H(y_);
    same_shape(x, y, "MultiInnerProdDiag");
    std::vector<int> pos = mask_positions(mask, n);
    BX_REQUIRE((int)pos.size() == x->num_active_vectors,
               "MultiInnerProdDiag: mask selects %d entries for %d active columns", (int)pos.size(),
               (int)x->num_active_vectors);
    if (pos.empty()) return;
    if (cplx(x)) {
        komplex *d = (komplex *)scratch(3, pos.size() * sizeof(komplex));
        zdiagdot_cols(x->size, x->data, x->ld, colmap_of(x), y->data, y->ld, colmap_of(y), d);
        std::vector<komplex> tmp(pos.size());
        download_sync(tmp.data(), d, tmp.size() * sizeof(komplex));
        for (size_t i = 0; i < pos.size(); i++) ((komplex *)diag)[pos[i]] = tmp[i];
        return;
    }
    double *d = (double *)scratch(3, pos.size() * sizeof(double));
    diagdot_cols(x->precision, x->size, x->data, x->ld, colmap_of(x), y->data, y->ld,
                 colmap_of(y), d);
    std::vector<double> tmp(pos.size());
    download_sync(tmp.data(), d, tmp.size() * sizeof(double));
    for (size_t i = 0; i < pos.size(); i++) ((double *)diag)[pos[i]] = tmp[i];
}

static void vecmat_common(void *x_, BlopexInt gh, BlopexInt h, BlopexInt w, void *v, void *y_,
                          bool acc) {
    hip_Multi_Vector *x = H(x_), *y = H(y_);
    same_shape(x, y, "MultiVecMat");
    BX_REQUIRE(h > 0, "MultiVecMat: empty coefficient matrix");
    BX_REQUIRE(h == x->num_active_vectors && w == y->num_active_vectors,
               "MultiVecMat: %d x %d matrix for %d -> %d active columns", (int)h, (int)w,
               (int)x->num_active_vectors, (int)y->num_active_vectors);
    BX_REQUIRE(x != y, "MultiVecMat: x and y must differ");
    if (cplx(x)) {
        std::vector<komplex> c((size_t)h * w);
        for (BlopexInt j = 0; j < w; j++)
            memcpy(c.data() + (size_t)j * h, (const komplex *)v + (size_t)j * gh,
                   sizeof(komplex) * h);
        void *dc = scratch(2, c.size() * sizeof(komplex));
        upload_async(dc, c.data(), c.size() * sizeof(komplex));
        zvecmat_cols(x->size, x->data, x->ld, colmap_of(x), dc, h, y->data, y->ld, colmap_of(y),
                     acc);
        return;
    }
    std::vector<double> c((size_t)h * w);
    for (BlopexInt j = 0; j < w; j++)
        memcpy(c.data() + (size_t)j * h, (const double *)v + (size_t)j * gh, sizeof(double) * h);
    double *dc = (double *)scratch(2, c.size() * sizeof(double));
    upload_async(dc, c.data(), c.size() * sizeof(double));
    vecmat_cols(x->precision, x->size, x->data, x->ld, colmap_of(x), dc, h, y->data, y->ld,
                colmap_of(y), acc);
}

static void hipmv_VecMat(void *x, BlopexInt gh, BlopexInt h, BlopexInt w, void *v, void *y) {
    vecmat_common(x, gh, h, w, v, y, false);
}

static void hipmv_Xapy(void *x, BlopexInt gh, BlopexInt h, BlopexInt w, void *v, void *y) {
    vecmat_common(x, gh, h, w, v, y, true);
}

static void hipmv_VecMatDiag(void *x_, BlopexInt *mask, BlopexInt n, void *diag, void *y_) {
    hip_Multi_Vector *x = H(x_), *y = H(y_);
    same_shape(x, y, "MultiVecMatDiag");
    std::vector<int> pos = mask_positions(mask, n);
    BX_REQUIRE((int)pos.size() == x->num_active_vectors,
               "MultiVecMatDiag: mask selects %d entries for %d active columns", (int)pos.size(),
               (int)x->num_active_vectors);
    if (cplx(x)) {
        if (pos.empty()) return;
        std::vector<komplex> coef(pos.size());
        for (size_t i = 0; i < pos.size(); i++) coef[i] = ((const komplex *)diag)[pos[i]];
        void *dc = scratch(2, coef.size() * sizeof(komplex));
        upload_async(dc, coef.data(), coef.size() * sizeof(komplex));
        zscale_cols(x->size, dc, x->data, x->ld, colmap_of(x), y->data, y->ld, colmap_of(y));
        return;
    }
    std::vector<double> coef(pos.size());
    for (size_t i = 0; i < pos.size(); i++) coef[i] = ((const double *)diag)[pos[i]];
    scale_cols(x->precision, x->size, coef.data(), x->data, x->ld, colmap_of(x), y->data,
               y->ld, colmap_of(y));
}

static void hipmv_Axpy(double alpha, void *x_, void *y_) {
    hip_Multi_Vector *x = H(x_), *y = H(y_);
    same_shape(x, y, "MultiAxpy");
    const View vx = view(x), vy = view(y);
    axpy_cols(vx.prec, vx.n, alpha, x->data, vx.ld, colmap_of(x), y->data, vy.ld, colmap_of(y));
}

static void hipmv_Print(void *x_, char *tag, BlopexInt limit) {
    hip_Multi_Vector *x = H(x_);
    const int c = cplx(x) ? 2 : 1;
    std::vector<double> host((size_t)x->size * x->num_vectors * c);
    hip_Multi_VectorDownload(x, host.data());
    printf("======= %s =========\n", tag ? tag : "");
    printf("size %ld\nowns data %d\nnum vectors %d\nnum active vectors  %d\n", x->size,
           (int)x->owns_data, (int)x->num_vectors, (int)x->num_active_vectors);
    long rows = x->size;
    int cols = x->num_active_vectors;
    if (limit) {
        if (rows > limit) rows = limit;
        if (cols > limit) cols = limit;
    }
    for (int i = 0; i < cols; i++) printf("index %d active %d\n", i, (int)x->active_indices[i]);
    for (int i = 0; i < cols; i++)
        for (long r = 0; r < rows; r++) {
            const size_t e = ((size_t)x->active_indices[i] * x->size + r) * c;
            if (c == 2)
                printf("%ld %d  %22.16e  %22.16e \n", r, i, host[e], host[e + 1]);
            else
                printf("%ld %d  %22.16e  \n", r, i, host[e]);
        }
}

static void hipmv_Eval(void (*f)(void *, void *, void *), void *par, void *x_, void *y_) {
    hip_Multi_Vector *x = H(x_), *y = H(y_);
    same_shape(x, y, "Eval");
    BX_REQUIRE(x->num_active_vectors == y->num_active_vectors, "Eval: active column counts differ");
    size_t es = elem_size(x->precision);
    for (int j = 0; j < x->num_active_vectors; j++) {
        hip_Multi_Vector *xc = hip_Multi_VectorWrapDevice(
            (char *)x->data + (size_t)x->active_indices[j] * x->ld * es, x->size, 1, x->precision);
        hip_Multi_Vector *yc = hip_Multi_VectorWrapDevice(
            (char *)y->data + (size_t)y->active_indices[j] * y->ld * es, y->size, 1, y->precision);
        f(par, xc, yc);
        hip_Multi_VectorDestroy(xc);
        hip_Multi_VectorDestroy(yc);
    }
}

BlopexInt HipSetupInterpreter(mv_InterfaceInterpreter *i) {
    memset(i, 0, sizeof(*i));
    i->CopyCreateMultiVector = hipmv_CopyCreate;
    i->DestroyMultiVector = hipmv_Destroy;
    i->Width = hipmv_Width;
    i->Height = hipmv_Height;
    i->SetMask = hipmv_SetMask;
    i->CopyMultiVector = hipmv_Copy;
    i->ClearMultiVector = hipmv_Clear;
    i->SetRandomVectors = hipmv_SetRandom;
    i->MultiInnerProd = hipmv_InnerProd;
    i->MultiInnerProdDiag = hipmv_InnerProdDiag;
    i->MultiVecMat = hipmv_VecMat;
    i->MultiVecMatDiag = hipmv_VecMatDiag;
    i->MultiAxpy = hipmv_Axpy;
    i->MultiPrint = hipmv_Print;
    i->MultiXapy = hipmv_Xapy;
    i->Eval = hipmv_Eval;
    return 0;
}

int blopex_amd_is_hip_interpreter(const mv_InterfaceInterpreter *i) {
    return i && i->CopyCreateMultiVector == hipmv_CopyCreate;
}

// ---- operators -------------------------------------------------------------

// Diagonal copy of A for the gather-free product (see hip_CSRMatrix): offsets
// c - r of all nonzeros; at most HIP_CSR_MAX_DIAGONALS of them, at least half
// of the diagonal slots filled, sorted distinct columns in every row.
// A 7-point stencil on an nx x ny x nz grid (x fastest): diagonals at
// -nx*ny, -nx, -1, 0, 1, nx, nx*ny with zero coefficients wherever the
// neighbour lies outside the grid (the -1 entry of an x = 0 row would
// otherwise couple to the end of the previous line).  dv: ndiag x diag_ld.
static void find_grid(hip_CSRMatrix *A, const double *dv) {
    A->grid[0] = A->grid[1] = A->grid[2] = 0;
    if (A->ndiag != 7) return;
    const int *o = A->diag_offset;
    const long lx = o[5], pl = o[6], n = A->n;
    if (o[2] != -1 || o[3] != 0 || o[4] != 1) return;
    if (o[1] != -lx || o[0] != -pl || o[6] != pl || lx < 2 || pl % lx || pl / lx < 2 || n % pl)
        return;
    const long ny = pl / lx, nz = n / pl, ld = A->diag_ld;
    for (long r = 0; r < n; r++) {
        const long x = r % lx, y = (r / lx) % ny;
        if ((x == 0 && dv[2 * ld + r] != 0.0) || (x == lx - 1 && dv[4 * ld + r] != 0.0) ||
            (y == 0 && dv[1 * ld + r] != 0.0) || (y == ny - 1 && dv[5 * ld + r] != 0.0))
            return;
    }
    A->grid[0] = (int)lx;
    A->grid[1] = (int)ny;
    A->grid[2] = (int)nz;
}

static void build_diagonals(hip_CSRMatrix *A, const int *rowptr, const int *colind,
                            const double *val) {
    const long n = A->n, nnz = A->nnz;
    A->ndiag = 0;
    if (n < 4 || nnz == 0) return;
    std::vector<long> offs;
    for (long r = 0; r < n; r++) {
        for (int e = rowptr[r]; e < rowptr[r + 1]; e++) {
            if (e > rowptr[r] && colind[e] <= colind[e - 1]) return;  // unsorted or repeated
            const long o = (long)colind[e] - r;
            if (std::find(offs.begin(), offs.end(), o) == offs.end()) {
                if ((int)offs.size() == HIP_CSR_MAX_DIAGONALS) return;
                offs.push_back(o);
            }
        }
    }
    if (2 * nnz < (long)offs.size() * n) return;  // too sparse for diagonals
    std::sort(offs.begin(), offs.end());
    const int nd = (int)offs.size();
    const long ld = (n + 3) / 4 * 4;
    std::vector<double> dv((size_t)nd * ld, 0.0);
    for (long r = 0; r < n; r++)
        for (int e = rowptr[r]; e < rowptr[r + 1]; e++) {
            const int d = (int)(std::lower_bound(offs.begin(), offs.end(), (long)colind[e] - r) -
                                offs.begin());
            dv[(size_t)d * ld + r] = val[e];
        }
    A->ndiag = nd;
    for (int d = 0; d < nd; d++) A->diag_offset[d] = (int)offs[d];
    A->diag_ld = ld;
    find_grid(A, dv.data());
    BX_HIP(hipMalloc(&A->diag_val, sizeof(double) * nd * ld));
    BX_HIP(hipMalloc(&A->diag_val32, sizeof(float) * nd * ld));
    BX_HIP(hipMemcpy(A->diag_val, dv.data(), sizeof(double) * nd * ld, hipMemcpyHostToDevice));
    convert_block((long)nd * ld, 1, A->diag_val, (long)nd * ld, HIPMV_FP64, A->diag_val32,
                  (long)nd * ld, HIPMV_FP32);
}

hip_CSRMatrix *hip_CSRMatrixCreate(long n, long nnz, const int *rowptr, const int *colind,
                                   const double *val) {
    BX_REQUIRE(n >= 0 && nnz >= 0, "hip_CSRMatrixCreate: bad sizes");
    BX_REQUIRE(rowptr[n] == nnz, "hip_CSRMatrixCreate: rowptr[n] = %d, nnz = %ld", rowptr[n], nnz);
    for (long r = 0; r < n; r++)
        BX_REQUIRE(rowptr[r] <= rowptr[r + 1], "hip_CSRMatrixCreate: rowptr decreases at row %ld", r);
    for (long e = 0; e < nnz; e++)
        BX_REQUIRE(colind[e] >= 0 && colind[e] < n, "hip_CSRMatrixCreate: column %d out of range",
                   colind[e]);
    hip_CSRMatrix *A = (hip_CSRMatrix *)calloc(1, sizeof(hip_CSRMatrix));
    A->n = n;
    A->nnz = nnz;
    hipStream_t s = stream();
    BX_HIP(hipMalloc(&A->rowptr, sizeof(int) * (n + 1)));
    BX_HIP(hipMalloc(&A->colind, sizeof(int) * (nnz > 0 ? nnz : 1)));
    BX_HIP(hipMalloc(&A->val, sizeof(double) * (nnz > 0 ? nnz : 1)));
    BX_HIP(hipMalloc(&A->val32, sizeof(float) * (nnz > 0 ? nnz : 1)));
    BX_HIP(hipMemcpyAsync(A->rowptr, rowptr, sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    BX_HIP(hipMemcpyAsync(A->colind, colind, sizeof(int) * nnz, hipMemcpyHostToDevice, s));
    BX_HIP(hipMemcpyAsync(A->val, val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
    convert_block(nnz, 1, A->val, nnz, HIPMV_FP64, A->val32, nnz, HIPMV_FP32);
    build_diagonals(A, rowptr, colind, val);
    BX_HIP(hipStreamSynchronize(s));
    return A;
}

void hip_CSRMatrixDestroy(hip_CSRMatrix *A) {
    if (!A) return;
    BX_HIP(hipStreamSynchronize(stream()));
    BX_HIP(hipFree(A->rowptr));
    BX_HIP(hipFree(A->colind));
    BX_HIP(hipFree(A->val));
    BX_HIP(hipFree(A->val32));
    if (A->ndiag) {
        BX_HIP(hipFree(A->diag_val));
        BX_HIP(hipFree(A->diag_val32));
    }
    free(A);
}

int hip_CSRMatrixProductKernel(const hip_CSRMatrix *A, int precision, BlopexInt ncols, long ld) {
    BX_REQUIRE(A && (precision == HIPMV_FP64 || precision == HIPMV_FP32) && ld >= A->n,
               "hip_CSRMatrixProductKernel: bad arguments");
    return spmm_kernel(precision, A, ld, (int)ncols);
}

void HipCSRMatMultiVec(void *A_, void *X_, void *Y_) {
    hip_CSRMatrix *A = (hip_CSRMatrix *)A_;
    hip_Multi_Vector *X = H(X_), *Y = H(Y_);
    BX_REQUIRE(A && X && Y, "HipCSRMatMultiVec: NULL argument");
    same_shape(X, Y, "HipCSRMatMultiVec");
    BX_REQUIRE(X->size == A->n, "HipCSRMatMultiVec: block has %ld rows, matrix %ld", X->size, A->n);
    real_only(X, "HipCSRMatMultiVec");
    spmm_csr(X->precision, A, X->data, X->ld, colmap_of(X), Y->data, Y->ld, colmap_of(Y));
}

hip_DiagMatrix *hip_DiagMatrixCreate(long n, const double *d) {
    hip_DiagMatrix *D = (hip_DiagMatrix *)calloc(1, sizeof(hip_DiagMatrix));
    D->n = n;
    hipStream_t s = stream();
    BX_HIP(hipMalloc(&D->d, sizeof(double) * (n > 0 ? n : 1)));
    BX_HIP(hipMalloc(&D->d32, sizeof(float) * (n > 0 ? n : 1)));
    BX_HIP(hipMemcpyAsync(D->d, d, sizeof(double) * n, hipMemcpyHostToDevice, s));
    convert_block(n, 1, D->d, n, HIPMV_FP64, D->d32, n, HIPMV_FP32);
    BX_HIP(hipStreamSynchronize(s));
    return D;
}

hip_DiagMatrix *hip_DiagMatrixCreateJacobi(const hip_CSRMatrix *A) {
    std::vector<int> rp(A->n + 1), ci(A->nnz);
    std::vector<double> v(A->nnz);
    hipStream_t s = stream();
    BX_HIP(hipMemcpyAsync(rp.data(), A->rowptr, sizeof(int) * (A->n + 1), hipMemcpyDeviceToHost, s));
    BX_HIP(hipMemcpyAsync(ci.data(), A->colind, sizeof(int) * A->nnz, hipMemcpyDeviceToHost, s));
    BX_HIP(hipMemcpyAsync(v.data(), A->val, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, s));
    BX_HIP(hipStreamSynchronize(s));
    std::vector<double> d(A->n, 0.0);
    for (long r = 0; r < A->n; r++) {
        double a = 0.0;
        for (int e = rp[r]; e < rp[r + 1]; e++)
            if (ci[e] == r) a = v[e];
        BX_REQUIRE(a != 0.0, "hip_DiagMatrixCreateJacobi: zero diagonal in row %ld", r);
        d[r] = 1.0 / a;
    }
    return hip_DiagMatrixCreate(A->n, d.data());
}

void hip_DiagMatrixDestroy(hip_DiagMatrix *D) {
    if (!D) return;
    BX_HIP(hipStreamSynchronize(stream()));
    BX_HIP(hipFree(D->d));
    BX_HIP(hipFree(D->d32));
    free(D);
}

void HipDiagMatMultiVec(void *D_, void *X_, void *Y_) {
    hip_DiagMatrix *D = (hip_DiagMatrix *)D_;
    hip_Multi_Vector *X = H(X_), *Y = H(Y_);
    BX_REQUIRE(D && X && Y, "HipDiagMatMultiVec: NULL argument");
    same_shape(X, Y, "HipDiagMatMultiVec");
    BX_REQUIRE(X->size == D->n, "HipDiagMatMultiVec: block has %ld rows, diagonal %ld", X->size, D->n);
    real_only(X, "HipDiagMatMultiVec");
    diag_apply(X->precision, D, X->data, X->ld, colmap_of(X), Y->data, Y->ld, colmap_of(Y));
}

void HipDenseMatMultiVec(void *A_, void *X_, void *Y_) {
    const hip_Multi_Vector *X = H(X_), *Y = H(Y_);
    BX_REQUIRE(X->size == Y->size, "HipDenseMatMultiVec: X has %ld rows, Y %ld", X->size, Y->size);
    BX_REQUIRE(X->precision == Y->precision, "HipDenseMatMultiVec: X and Y differ in precision");
    if (!A_) {
        BX_REQUIRE(X->precision != HIPMV_C64, "HipDenseMatMultiVec: the 1D Laplacian takes real blocks");
        lapl1d_apply(X->precision, X->size, X->data, X->ld, colmap_of(X), Y->data, Y->ld,
                     colmap_of(Y));
        return;
    }
    const hip_Multi_Vector *A = H(A_);
    BX_REQUIRE(A->size == A->num_vectors && A->size == X->size && X->size == Y->size,
               "HipDenseMatMultiVec: A is %ld x %d, X has %ld rows, Y %ld", A->size,
               (int)A->num_vectors, X->size, Y->size);
    if (A->precision == HIPMV_C64) {
        BX_REQUIRE(X->precision == HIPMV_C64,
                   "HipDenseMatMultiVec: a complex A takes complex blocks");
        zdense_apply(A->size, A->data, A->ld, X->data, X->ld, colmap_of(X), Y->data, Y->ld,
                     colmap_of(Y));
        return;
    }
    BX_REQUIRE(A->precision == HIPMV_FP64 && X->precision != HIPMV_C64,
               "HipDenseMatMultiVec: a real A must be FP64 and take real blocks");
    dense_apply(X->precision, A->size, (const double *)A->data, A->ld, X->data, X->ld,
                colmap_of(X), Y->data, Y->ld, colmap_of(Y));
}

}  // extern "C"
