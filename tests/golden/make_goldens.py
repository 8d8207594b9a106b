"""Writes reference_goldens.json: inputs and expected outputs transcribed from the
reference's own test files (RadialBasisFunctions.jl v0.7.0). Each entry names the
test file and lines it comes from. The values are the closed-form expectations the
reference tests assert; nothing here is computed by the reference."""
import json
import math
import pathlib

s5 = math.sqrt(5.0)
s21 = math.sqrt(21.0)
G = {
    "basis_derivatives": {
        "x1": [1.0, 2.0],
        "x2": [2.0, 4.0],
        "dim": 1,
        "cases": [
            {"basis": "PHS1", "d1": -1 / s5, "grad": [-1 / s5, -2 / s5], "d2": 4 / (5 * s5),
             "src": "test/basis/polyharmonic_spline.jl:43-45"},
            {"basis": "PHS3", "d1": -3 * s5, "grad": [-3 * s5, -6 * s5], "d2": 18 / s5,
             "src": "test/basis/polyharmonic_spline.jl:110-112"},
            {"basis": "PHS5", "d1": -25 * s5, "grad": [-25 * s5, -50 * s5], "d2": 40 * s5,
             "src": "test/basis/polyharmonic_spline.jl:177-179"},
            {"basis": "PHS7", "d1": -175 * s5, "grad": [-175 * s5, -350 * s5], "d2": 350 * s5,
             "src": "test/basis/polyharmonic_spline.jl:243-245"},
            {"basis": "IMQ", "eps": 2.0, "d1": 4 / (21 * s21), "grad": [4 / (21 * s21), 8 / (21 * s21)],
             "d2": -4 / (49 * s21), "src": "test/basis/inverse_multiquadric.jl:40-42"},
            {"basis": "Gaussian", "eps": 2.0, "d1": 8 / math.exp(20), "grad": [8 / math.exp(20), 16 / math.exp(20)],
             "d2": 56 / math.exp(20), "src": "test/basis/gaussian.jl:40-42"},
        ],
    },
    # phi(x1, x2) at x1 = (1, 2), x2 = (2, 4), r = sqrt(5): the closed forms the
    # reference's "Distances" testsets assert
    "basis_values": {
        "x1": [1.0, 2.0],
        "x2": [2.0, 4.0],
        "cases": [
            {"basis": "PHS1", "value": s5, "src": "test/basis/polyharmonic_spline.jl:33-35"},
            {"basis": "PHS3", "value": 5 * s5, "src": "test/basis/polyharmonic_spline.jl:100-102"},
            {"basis": "PHS5", "value": 25 * s5, "src": "test/basis/polyharmonic_spline.jl:167-169"},
            {"basis": "PHS7", "value": 125 * s5, "src": "test/basis/polyharmonic_spline.jl:233-235"},
            {"basis": "IMQ", "eps": 2.0, "value": 1 / s21, "src": "test/basis/inverse_multiquadric.jl:29-31"},
            {"basis": "Gaussian", "eps": 2.0, "value": math.exp(-20), "src": "test/basis/gaussian.jl:29-31"},
        ],
    },
    "monomials": [
        {"dim": 1, "deg": 0, "x": [2.0], "val": [1], "d1": [[0]], "d2": [[0]], "lap": [0],
         "src": "test/basis/monomial.jl:11-26"},
        {"dim": 1, "deg": 1, "x": [2.0], "val": [1, 2], "d1": [[0, 1]], "d2": [[0, 0]], "lap": [0, 0],
         "src": "test/basis/monomial.jl:28-49"},
        {"dim": 1, "deg": 2, "x": [2.0], "val": [1, 2, 4], "d1": [[0, 1, 4]], "d2": [[0, 0, 2]], "lap": [0, 0, 2],
         "src": "test/basis/monomial.jl:51-67"},
        {"dim": 2, "deg": 0, "x": [2.0, 3.0], "val": [1], "d1": [[0], [0]], "d2": [[0], [0]], "lap": [0],
         "src": "test/basis/monomial.jl:69-85"},
        {"dim": 2, "deg": 1, "x": [2.0, 3.0], "val": [1, 2, 3], "d1": [[0, 1, 0], [0, 0, 1]],
         "d2": [[0, 0, 0], [0, 0, 0]], "lap": [0, 0, 0], "src": "test/basis/monomial.jl:87-103"},
        {"dim": 2, "deg": 2, "x": [2.0, 3.0], "val": [1, 2, 3, 6, 4, 9],
         "d1": [[0, 1, 0, 3, 4, 0], [0, 0, 1, 2, 0, 6]], "d2": [[0, 0, 0, 0, 2, 0], [0, 0, 0, 0, 0, 2]],
         "lap": [0, 0, 0, 0, 2, 2], "src": "test/basis/monomial.jl:105-121"},
        {"dim": 2, "deg": 3, "x": [2.0, 3.0], "val": [8, 12, 4, 18, 6, 2, 27, 9, 3, 1],
         "d1": [[12, 12, 4, 9, 3, 1, 0, 0, 0, 0], [0, 4, 0, 12, 2, 0, 27, 6, 1, 0]],
         "d2": [[12, 6, 2, 0, 0, 0, 0, 0, 0, 0], [0, 0, 0, 4, 0, 0, 18, 2, 0, 0]],
         "lap": [12, 6, 2, 4, 0, 0, 18, 2, 0, 0], "src": "test/basis/monomial.jl:123-139"},
        {"dim": 3, "deg": 1, "x": [2.0, 3.0, 4.0], "val": [1, 2, 3, 4],
         "d1": [[0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], "d2": [[0, 0, 0, 0], [0, 0, 0, 0], [0, 0, 0, 0]],
         "lap": [0, 0, 0, 0], "src": "test/basis/monomial.jl:161-179"},
        {"dim": 3, "deg": 2, "x": [2.0, 3.0, 4.0], "val": [1, 2, 3, 4, 6, 8, 12, 4, 9, 16],
         "d1": [[0, 1, 0, 0, 3, 4, 0, 4, 0, 0], [0, 0, 1, 0, 2, 0, 4, 0, 6, 0], [0, 0, 0, 1, 0, 2, 3, 0, 0, 8]],
         "d2": [[0, 0, 0, 0, 0, 0, 0, 2, 0, 0], [0, 0, 0, 0, 0, 0, 0, 0, 2, 0], [0, 0, 0, 0, 0, 0, 0, 0, 0, 2]],
         "lap": [0, 0, 0, 0, 0, 0, 0, 2, 2, 2], "src": "test/basis/monomial.jl:181-199"},
        {"dim": 4, "deg": 0, "x": [2.0, 3.0, 4.0, 5.0], "val": [1], "d1": [[0], [0], [0], [0]],
         "d2": [[0], [0], [0], [0]], "lap": [0], "src": "test/basis/monomial.jl:201-217"},
    ],
    "hermite_poly_entries_1d": {
        "src": "test/solve/hermite_stencil.jl:108-158",
        "mon": [1, 1],
        "points": [[0.0], [0.5], [1.0]],
        "cases": [
            {"point": 0, "alpha": 0.0, "beta": 0.0, "boundary": False, "normal": [0.0], "expected": [1.0, 0.0]},
            {"point": 1, "alpha": 1.0, "beta": 0.0, "boundary": True, "normal": [0.0], "expected": [1.0, 0.5]},
            {"point": 2, "alpha": 0.0, "beta": 1.0, "boundary": True, "normal": [1.0], "expected": [0.0, 1.0]},
            {"point": 2, "alpha": 0.5, "beta": 0.5, "boundary": True, "normal": [1.0], "expected": [0.5, 1.0]},
        ],
    },
    "autoselect_k": {
        "src": "test/utils.jl:45-53",
        "points": [[0.01 * i, 0.02 * i * i] for i in range(1, 51)],
        "cases": [{"poly_deg": 2, "n": 50, "k": 12}, {"poly_deg": 1, "n": 50, "k": 6},
                  {"poly_deg": 0, "n": 50, "k": 5}, {"poly_deg": 2, "n": 5, "k": 5}],
    },
    "knn_small": {
        "src": "test/utils.jl:28-43",
        "points": [[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [2.0, 2.0]],
        "k": 2,
        "self_first": True,
        "second_of_0_in": [1, 2],
        "farthest_excluded": {"query": 3, "excluded": 0},
        "eval": {"points": [[0.9, 0.1]], "k": 3, "first": 1},
    },
    "hermite_domain_12": {
        "src": "test/hermite.jl:25-48 (geometry), :50-140 (exact polynomial reproduction)",
        "points": [[0.0, 0.0], [0.15, 0.1], [0.2, 0.25], [0.3, 0.15], [0.4, 0.3], [0.5, 0.2], [0.6, 0.35],
                   [0.7, 0.25], [0.75, 0.15], [0.85, 0.3], [0.9, 0.2], [1.0, 0.0]],
        "is_boundary": [True] + [False] * 10 + [True],
        "normals": [[1.0, 0.0], [-1.0, 0.0]],
        "cases": [
            {"bcs": [[1, 0], [1, 0]], "field": "x2+y2", "interior_laplacian": 4.0, "tol": 1e-10},
            {"bcs": [[0, 1], [0, 1]], "field": "y2", "interior_laplacian": 2.0, "tol": 1e-10},
            {"bcs": [[1, 0], [0, 1]], "field": "y2", "interior_laplacian": 2.0, "tol": 1e-10},
            {"bcs": [[1, 1], [2, 3]], "field": "y2", "interior_laplacian": 2.0, "tol": 1e-10},
        ],
    },
    "regrid_linear": {
        "src": "test/utils.jl:165-181",
        "targets": [[0.3, 0.2], [0.6, 0.25], [0.8, 0.2]],
        "field": [1.0, 2.0, -0.5],
        "atol": 1e-8,
    },
    "collocation_small": {
        "src": "test/solve.jl:7-47",
        "points": [[1.0, 2.0], [2.0, 1.0], [1.5, 0.0]],
        "eval_point": [0.0, 0.0],
    },
}

if __name__ == "__main__":
    out = pathlib.Path(__file__).with_name("reference_goldens.json")
    out.write_text(json.dumps(G, indent=1))
    print("wrote", out)
