"""The C++ API as an MRCPP caller sees it: tests/cpp/mrcpp_api_poisson.cpp
includes "MRCPP/MWFunctions", "MRCPP/MWOperators", "MRCPP/Gaussians",
"MRCPP/Printer" and "MRCPP/Timer", uses FunctionTree<3>, FunctionTreeVector<3>,
BoundingBox<3>(scale, corner, boxes), GaussFunc<3>(beta, alpha, pos, power),
PoissonOperator(mra, prec), build_grid, project, apply (plain, precTrees,
CompFunction), add, multiply and dot with MRCPP's signatures, and is built by
g++ against libmrcpp_amd.so (tests/cpp/Makefile, run by the build step).

Without a GPU the program must link, load, pass its host-side checks and stop
with the library's device error; on the MI355X every check must hold
(Coulomb self-energy within 10 prec of sqrt(2 beta / pi), the potential
erf(sqrt(beta) r)/r at r = 1, the overloads' results)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "bin", "mrcpp_api_poisson")


def _run():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/bin/mrcpp_api_poisson is missing: run the build step (make -C tests/cpp)")
    # a child process of its own (the program initialises its own device)
    return subprocess.run([EXE], capture_output=True, text=True, timeout=600)


def test_cpp_api_program_links_and_runs_host_part(mr):
    r = _run()
    print(r.stdout)
    assert "ok   Coulomb self-energy closed form" in r.stdout
    if mr.lib().mwg_device_count() < 1:
        assert r.returncode == 2 and "ERROR:" in r.stdout
    else:
        assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_api_program_on_gpu():
    r = _run()
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED: 0 failed checks" in r.stdout
    assert "FAIL " not in r.stdout
