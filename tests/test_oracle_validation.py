"""The reference's exact-solution validation cases, solved by the FP64 oracle
(CPU): pins the oracle's SIMPLE iteration as a whole against closed forms
(tests/ValidationTests.cpp:98-431 of sidhanth-sridhar/CFD-SIMULATOR)."""
import numpy as np
import pytest

from helpers import check_validation, converge, validation_case

O = pytest.importorskip("oracle_py")

CASES = [("uniform", 3000, 1e-12), ("couette", 3000, 1e-12), ("poiseuille", 4000, 1e-9),
         pytest.param("blasius", 4000, 1e-8, marks=pytest.mark.slow)]


@pytest.mark.parametrize("name,max_it,tol", CASES)
def test_oracle_validation_case(name, max_it, tol):
    mesh, kind, vel, pres, field, st, exact = validation_case(O, name)
    s = O.Solver(mesh, kind, vel, pres, st)
    s.initialise(*field)
    mon, taken = converge(s, max_it, tol)
    u, v, p, _ = s.field()
    check_validation(name, mesh, u, v, p, mon, s.flux(), exact, taken, max_it)


def test_oracle_residuals_fall_and_mass_is_conserved():
    """ValidationTests.cpp:383-431: 400 iterations, continuity / 1e6, boundary net flux 0."""
    mesh, kind, vel, pres, field, st, exact = validation_case(O, "mass_balance")
    s = O.Solver(mesh, kind, vel, pres, st)
    s.initialise(*field)
    cont = [s.iterate(1).continuity for _ in range(400)]
    assert cont[-1] < 1e-6 * cont[0]
    check_validation("mass_balance", mesh, *s.field()[:3], None, s.flux(), exact, 400, 400)
