"""ctypes mirror of include/rbfx.h (the library's C ABI).

Every struct here must match the header byte for byte; tests compare the
ctypes sizes against sizeof() reported by compiled C.
"""
import ctypes as C

# error codes
RBFX_OK = 0
RBFX_EINVAL = 1
RBFX_EDIM = 2
RBFX_EDEVICE = 3
RBFX_ESINGULAR = 4
RBFX_EUNSUPPORTED = 5

# basis kinds
RBFX_PHS1 = 1
RBFX_PHS3 = 3
RBFX_PHS5 = 5
RBFX_PHS7 = 7
RBFX_IMQ = 10
RBFX_GAUSSIAN = 11

# operator term kinds
OP_VALUE = 0
OP_D1 = 1
OP_D2 = 2
OP_MIXED = 3
OP_LAP = 4

# precision modes
PREC_FP64 = 0
PREC_FP32_REFINED = 1
PREC_LEARNED = 2

# per-stencil status
ST_OK = 0
ST_DIRICHLET = 1
ST_REFINED = 2
ST_FALLBACK64 = 3
ST_LEARNED = 4
ST_FP64 = 5
ST_SINGULAR = 8
ST_EVAL_NOT_FOUND = 9
ST_TOO_LARGE = 10

MAX_TERMS = 16
MAX_COLS = 8
MAX_N = 64
ABI_VERSION = 3

# surrogate network shape (must match csrc/mlp.h)
MLP_LAYERS = 4
MLP_IN_PAD = 128
MLP_HID = 256
MLP_OUT_PAD = 32


def mlp_k(l):
    return MLP_IN_PAD if l == 0 else MLP_HID


def mlp_n(l):
    return MLP_OUT_PAD if l == MLP_LAYERS - 1 else MLP_HID


def mlp_w_off(l):
    """offset of Wt_l [N_l x K_l] in the flat FP32 parameter array (b_l follows it)"""
    return sum(mlp_n(i) * mlp_k(i) + mlp_n(i) for i in range(l))


def mlp_b_off(l):
    return mlp_w_off(l) + mlp_n(l) * mlp_k(l)


def mlp_wb_off(l):
    """offset of Wt_l in the BF16 weight copy"""
    return sum(mlp_n(i) * mlp_k(i) for i in range(l))


MLP_N_PARAMS = mlp_w_off(MLP_LAYERS)
MLP_N_WBF16 = mlp_wb_off(MLP_LAYERS)


class Basis(C.Structure):
    _fields_ = [("kind", C.c_int32), ("poly_deg", C.c_int32), ("eps", C.c_double)]


class Term(C.Structure):
    _fields_ = [
        ("col", C.c_int32),
        ("kind", C.c_int32),
        ("dim1", C.c_int32),
        ("dim2", C.c_int32),
        ("coeff", C.c_double),
    ]


class Operator(C.Structure):
    _fields_ = [("n_cols", C.c_int32), ("n_terms", C.c_int32), ("terms", Term * MAX_TERMS)]


class BuildArgs(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("k", C.c_int32),
        ("n_data", C.c_int64),
        ("n_eval", C.c_int64),
        ("data", C.c_void_p),
        ("eval", C.c_void_p),
        ("adj_ptr", C.c_void_p),
        ("adj_idx", C.c_void_p),
        ("basis", Basis),
        ("op", Operator),
        ("is_boundary", C.c_void_p),
        ("bnd_index", C.c_void_p),
        ("bc_alpha", C.c_void_p),
        ("bc_beta", C.c_void_p),
        ("normals", C.c_void_p),
        ("n_bnd", C.c_int64),
        ("out_ptr", C.c_void_p),
        ("out_col", C.c_void_p),
        ("out_val", C.c_void_p),
        ("out_nnz", C.c_int64),
        ("status", C.c_void_p),
        ("precision", C.c_int32),
        ("refine_steps", C.c_int32),
        ("accept_tol", C.c_double),
        ("stream", C.c_void_p),
        ("out_lambda", C.c_void_p),
        ("model", C.c_void_p),
        ("poly_deg_map", C.c_void_p),
        ("order", C.c_void_p),
    ]


class MlpModel(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("k", C.c_int32),
        ("basis_kind", C.c_int32),
        ("poly_deg", C.c_int32),
        ("op_kind", C.c_int32),
        ("n_out", C.c_int32),
        ("params", C.c_void_p),
        ("wbf16", C.c_void_p),
        ("out_mean", C.c_void_p),
        ("out_std", C.c_void_p),
        ("accept_tol", C.c_double),
        ("out_rho", C.c_void_p),
    ]


class MlpRunArgs(C.Structure):
    _fields_ = [
        ("batch", C.c_int32),
        ("n_out", C.c_int32),
        ("x", C.c_void_p),
        ("target", C.c_void_p),
        ("params", C.c_void_p),
        ("wbf16", C.c_void_p),
        ("grads", C.c_void_p),
        ("loss", C.c_void_p),
        ("y_out", C.c_void_p),
        ("workspace", C.c_void_p),
        ("stream", C.c_void_p),
        ("row_weight", C.c_void_p),
    ]


class SpmvArgs(C.Structure):
    _fields_ = [
        ("n_rows", C.c_int64),
        ("n_cols", C.c_int64),
        ("row_ptr", C.c_void_p),
        ("k", C.c_int32),
        ("col_idx", C.c_void_p),
        ("vals", C.c_void_p),
        ("nnz", C.c_int64),
        ("n_valarrays", C.c_int32),
        ("n_fields", C.c_int32),
        ("x", C.c_void_p),
        ("x_ld_point", C.c_int64),
        ("x_ld_field", C.c_int64),
        ("y", C.c_void_p),
        ("y_ld_point", C.c_int64),
        ("y_ld_field", C.c_int64),
        ("y_ld_val", C.c_int64),
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("stream", C.c_void_p),
        ("row_order", C.c_void_p),
        ("rows_permuted", C.c_int32),
        ("max_row_len", C.c_int64),
        ("mapping", C.c_int32),
        ("tile_rows", C.c_int32),
        ("tile_src", C.c_void_p),
        ("tile_src_off", C.c_void_p),
        ("tile_src_n", C.c_void_p),
        ("tile_col", C.c_void_p),
    ]


def make_operator(n_cols, terms):
    """terms: iterable of (col, kind, dim1, dim2, coeff) with 0-based dims."""
    terms = list(terms)
    if len(terms) > MAX_TERMS:
        raise ValueError(f"operator has {len(terms)} terms; the ABI carries at most {MAX_TERMS}")
    if not 1 <= n_cols <= MAX_COLS:
        raise ValueError(f"operator has {n_cols} columns; the ABI carries at most {MAX_COLS}")
    op = Operator()
    op.n_cols = n_cols
    op.n_terms = len(terms)
    for i, (col, kind, d1, d2, coeff) in enumerate(terms):
        op.terms[i] = Term(col, kind, d1, d2, float(coeff))
    return op
