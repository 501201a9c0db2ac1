"""The oracle's procedural equation layer (oracle/equ.py) against the
reference's own test expectations (tests/equ_cases.py: procedural_equation_test,
public_equation_test, physics_contract_test) and the definitions of the
reductions.  The same scenarios run on the HIP library in test_gpu_equ.py."""
import types

import equ as oracle_equ
import equ_cases as ec
import oracle_lib as orc

B = types.SimpleNamespace(**{k: getattr(oracle_equ, k) for k in dir(oracle_equ) if not k.startswith("_")})
B.mesh = orc.oracle_mesh


def test_procedural_equation_known_answers():
    ec.procedural(B)


def test_public_equation_known_answers():
    ec.public(B)


def test_frozen_diffusion_flux_belongs_to_the_matrix():
    ec.frozen_diffusion(B)


def test_courant_contract():
    ec.courant_contract(B)


def test_reductions():
    ec.reductions(B)
