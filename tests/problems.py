"""Problem descriptions shared by the tests (dicts with the tdz_problem fields)."""

# config 2 shape at reduced size: soft-core atom, quadratic CAP, z field
SOFTCORE_3D = dict(dim=3, nelem=(6, 5, 7), lo=(-8.0, -7.0, -9.0), hi=(8.0, 7.0, 9.0), p=4, nquad=8,
                   mass=1.0, charge=-1.0, pot_kind=2, pot=(1.0, 2.0), cap_kind=1, cap=(0.01, 3.0))
# uneven orders and a Manolopoulos CAP with a heavier carrier
MANO_3D = dict(dim=3, nelem=(4, 6, 5), lo=(-6.0, -6.0, -6.0), hi=(6.0, 6.0, 6.0), p=3, nquad=5,
               mass=1.7, charge=-1.0, pot_kind=2, pot=(1.0, 1.0), cap_kind=2, cap=(1.2,))
# config 4 shape: 2D soft-core, p = 5
SOFTCORE_2D = dict(dim=2, nelem=(12, 10), lo=(-10.0, -10.0), hi=(10.0, 10.0), p=5, nquad=8,
                   mass=1.0, charge=-1.0, pot_kind=2, pot=(1.0, 1.0), cap_kind=1, cap=(0.02, 3.0))
# config 1 shape: 1D harmonic oscillator with Dirichlet walls
HO_1D = dict(dim=1, nelem=(200,), lo=(-20.0,), hi=(20.0,), p=5, nquad=8, dirichlet=1,
             mass=1.0, charge=-1.0, pot_kind=1, pot=(1.0,))
# hole carrier (charge +1) inverts potential and coupling
HOLE_3D = dict(SOFTCORE_3D, charge=1.0, nelem=(4, 4, 4))
