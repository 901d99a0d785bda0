"""Mutation tests of the oracle's pins: each case builds oracle/frg_oracle.c with one deliberate,
plausible mistake planted (-DORACLE_MUTANT=k: a swapped operand, a dropped term, a wrong sign or
index; the sites are marked MUT(k) in the source) into a temporary directory, and shows that
the named check of tests/test_oracle.py fails on that build while it passes on the real one.
The numpy oracle of the step-size model (oracle/mlp_oracle.py) gets the same treatment from an
altered copy of its source executed as a separate module.  The mistakes live only in these
temporary CPU copies; nothing here touches a GPU."""
import concurrent.futures as cf
import functools
import os
import re
import subprocess
import types

import pytest

from oracle import oracle as O
from tests import test_mlp_oracle as TM
from tests import test_oracle as T

SRC = os.path.join(os.path.dirname(O.__file__), "frg_oracle.c")
INC = os.path.join(os.path.dirname(O.__file__), "fsys_expanded.inc")
CFLAGS = ["-std=gnu11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-O2"]

# (k, the mistake, the check that must catch it)
MUTANTS = [
    (1, "Philox: multipliers applied to swapped counter words", T.test_philox_known_answers),
    (2, "uniform draw: 32-bit halves of the 64-bit word swapped", T.test_uniform_draws),
    (3, "JVAR: wrong sign of the product-rule term of an unconjugated variable",
     T.test_jvar_jpar_finite_differences),
    (4, "JVAR: last term of K_j of equation I dropped", T.test_jvar_jpar_finite_differences),
    (5, "JPAR: columns Pc0 and Pc_j of equation II transposed", T.test_jvar_jpar_finite_differences),
    (6, "q(t): 1 - t and t transposed in the denominator", T.test_q_and_dqdt),
    (7, "SwitchPatch: t rescaled with the other coordinates", T.test_switch_patch_with_reference_patches),
    (8, "LinearSolve: rows exchanged without their right-hand sides", T.test_linear_solve_vs_lapack),
    (9, "Newton: update applied with the wrong sign", T.test_newton_converges_quadratically),
    (10, "RK6(5)8M: two weights of the propagated solution exchanged", T.test_rk65_orders),
    (11, "step-size factor with the exponent of a fifth-order estimate",
     T.test_step_factor_targets_the_tolerance),
    (12, "Cognate12: P - F taken the wrong way round", T.test_cognate_is_a_root),
    (13, "Lambert W: Halley step added instead of subtracted", T.test_lambertw),
    (14, "duplicate test: last variable not compared", T.test_dedupe_compares_every_real),
    (15, "GenerateStart: Qc set to Q instead of 1/Q", functools.partial(T.test_start_points_are_roots, 0)),
    (16, "LinearSolve: no row exchange (pivot search inverted)", T.test_linear_solve_vs_lapack),
    (17, "complex division: wrong sign in the imaginary part (Smith's first branch)",
     T.test_complex_division_and_sqrt),
    (18, "complex sqrt: sign of the imaginary part taken from the wrong half plane",
     T.test_complex_division_and_sqrt),
    (19, "IVP: arclength tangent not normalised", T.test_ivp_tangent_properties),
    (20, "LinearSolveNewton: x_t pinned to -dt", T.test_linear_solve_newton_ignores_last_row_and_pins_t),
    (21, "FSYS II_j: Fc read as F in the fourth term", T.test_fsys_matches_independent_form),
]


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")

    def build(k):
        out = str(d / f"liboracle_mutant{k}.so")
        r = subprocess.run(["gcc", *CFLAGS, f"-DORACLE_MUTANT={k}", "-o", out, SRC, "-lm"], capture_output=True,
                           text=True)
        assert r.returncode == 0, r.stderr
        return k, out

    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        return dict(ex.map(build, [k for k, _, _ in MUTANTS]))


def test_every_site_in_the_source_has_a_case():
    """Every MUT(k) site is exercised: here on the CPU, or (TrackPath's bookkeeping) by the
    comparison with the GPU tracker in tests/test_gpu_oracle_mutants.py."""
    from tests.test_gpu_oracle_mutants import TRACKPATH_MUTANTS
    src = open(SRC).read() + open(INC).read()
    sites = {int(x) for x in re.findall(r"MUT\((\d+)\)", src)}
    cpu, gpu = {k for k, _, _ in MUTANTS}, {k for k, _, _ in TRACKPATH_MUTANTS}
    assert not cpu & gpu and sites == cpu | gpu
    makefile = open(os.path.join(os.path.dirname(SRC), "Makefile")).read()
    assert re.search(r"TRACKPATH_MUTANTS = (.*)", makefile).group(1).split() == [str(k) for k in sorted(gpu)]


@pytest.mark.parametrize("k,what,check", MUTANTS, ids=[f"m{k}" for k, _, _ in MUTANTS])
def test_mutant_is_caught(mutant_libs, k, what, check):
    check()  # the real oracle passes
    with O.using(O.load(mutant_libs[k])):
        with pytest.raises(AssertionError):
            check()


MLP_SRC = os.path.join(os.path.dirname(O.__file__), "mlp_oracle.py")

# (the mistake, source text, altered text, the check that must catch it)
MLP_MUTANTS = [
    ("second layer with W2 transposed (forward and loss alike)", "h1 @ w2.T", "h1 @ w2",
     TM.check_forward_matches_scalar_definition),
    ("tanh derivative of the second layer dropped from the gradient", "np.outer(dy, w3) * (1 - h2 * h2)",
     "np.outer(dy, w3)", TM.check_gradient_finite_differences),
    ("gradient of W1 stored transposed", "(d1.T @ x).ravel()", "(x.T @ d1).ravel()",
     TM.check_gradient_finite_differences),
    ("features multiplied by their deviation in the forward pass",
     "x = (np.asarray(feats, np.float64) - xmu) / xsd", "x = (np.asarray(feats, np.float64) - xmu) * xsd",
     TM.check_forward_consistent_with_loss),
    ("output mean left out of the forward pass", "return y * ysd + ymu", "return y * ysd",
     TM.check_forward_consistent_with_loss),
]


def mlp_mutant(old, new):
    src = open(MLP_SRC).read()
    assert old in src, old
    mod = types.ModuleType("mlp_oracle_mutant")
    exec(compile(src.replace(old, new), MLP_SRC + " (mutant)", "exec"), mod.__dict__)
    return mod


@pytest.mark.parametrize("what,old,new,check", MLP_MUTANTS, ids=[f"mlp{i}" for i in range(len(MLP_MUTANTS))])
def test_mlp_mutant_is_caught(what, old, new, check):
    check()  # the real oracle passes
    with pytest.raises(AssertionError):
        check(mlp_mutant(old, new))
