"""The oracle's mass and CSV under the reference's subdomain split and MPI rank count
(hydro2d/main.cxx:166-265, 620-657; hydro3d/main.cxx:172-281): subdomain s = (a*sdy+b)*sdz+c
of size NB/sd (the last along an axis adding NB % sd), rank r owning s % ranks == r.

Pinned three ways: an X-only split on one rank walks the cells in the whole-domain order,
so it must equal the plain mass and CSV bit for bit and byte for byte; a Python
restatement of the split (80-bit long double, the reference's compensated sum) must give
the same mass for every split and rank count; and the CSV must hold every cell once, in
the restated block order."""
import numpy as np
import pytest

from gpu_helpers import random_state_2d, random_state_3d


def _boxes(nb, sd):
    """Per axis: list of (base, n) of the subdomains along it."""
    out = []
    for n, s in zip(nb, sd):
        size = n // s
        out.append([(size * a, size + (n % s if a == s - 1 else 0)) for a in range(s)])
    return out


def _order(nb, sd, ranks):
    """[(rank, box)] in file order: ranks in order, each rank's subdomains in index order."""
    dim = len(nb)
    sd = list(sd) + [1] * (3 - dim)
    bx = _boxes(list(nb) + [1] * (3 - dim), sd)
    subs = []
    for a in range(sd[0]):
        for b in range(sd[1]):
            for c in range(sd[2]):
                subs.append((bx[0][a], bx[1][b], bx[2][c]))
    return [(r, subs[s]) for r in range(ranks) for s in range(r, len(subs), ranks)]


def _mass_restated(rho, nb, sd, ranks):
    """rho: interior density array; per rank one compensated np.longdouble sum."""
    dim = len(nb)
    total = 0.0
    per = {}
    for r, box in _order(nb, sd, ranks):
        tm, cc = per.get(r, (np.longdouble(0), np.longdouble(0)))
        (i0, ni), (j0, nj), (k0, nk) = box
        for i in range(i0, i0 + ni):
            for j in range(j0, j0 + nj):
                for k in range(k0, k0 + nk):
                    var = np.longdouble(rho[i, j] if dim == 2 else rho[i, j, k])
                    t = tm + var
                    if abs(tm) >= var:
                        cc += (tm - t) + var
                    else:
                        cc += (var - t) + tm
                    tm = t
        per[r] = (tm, cc)
    for r in range(ranks):
        if r in per:
            total += float(per[r][0] + per[r][1])
    return total


def _state(dim, nb, seed):
    if dim == 2:
        return random_state_2d(*nb, seed=seed)
    return random_state_3d(*nb, seed=seed)


@pytest.mark.parametrize("dim,nb", [(2, (37, 23)), (3, (9, 8, 7))])
def test_x_split_on_one_rank_is_the_plain_order(oracle, tmp_path, dim, nb):
    rhoE, mom = _state(dim, nb, seed=11)
    plain = oracle.mass2d(rhoE) if dim == 2 else oracle.mass3d(rhoE)
    for sdx in (1, 2, 3, 5):
        sd = (sdx,) + (1,) * (dim - 1)
        assert oracle.mass_sd(rhoE, sd, 1) == plain
    want = tmp_path / "plain.csv"
    (oracle.write_csv2d if dim == 2 else oracle.write_csv3d)(want, rhoE, mom)
    got = tmp_path / "sd.csv"
    oracle.write_csv_sd(got, rhoE, mom, (3,) + (1,) * (dim - 1), 1)
    assert got.read_bytes() == want.read_bytes()


@pytest.mark.parametrize("dim,nb,sd", [(2, (13, 11), (2, 3)), (2, (17, 9), (4, 2)), (3, (7, 6, 5), (2, 2, 2)),
                                       (3, (9, 5, 6), (3, 1, 2))])
@pytest.mark.parametrize("ranks", [1, 2, 3, 4])
def test_mass_under_split_and_ranks(oracle, dim, nb, sd, ranks):
    """Small integer densities around two cells of 2^60 and 2^59: a rank's long double
    sum is exact, but its rounding to double (steps of 256 at 2^60) drops the small
    cells it holds, so the total depends on which cells each rank sums.  The rank
    counts and splits give several different totals, and each must be the restated one."""
    rhoE, _ = _state(dim, nb, seed=7 + ranks)
    rng = np.random.default_rng(ranks)
    interior = (slice(1, -1),) * dim
    rho = rng.integers(1, 4, size=nb).astype(np.float64)
    rho[(0,) * dim] = 2.0**60
    rho[tuple(n - 1 for n in nb)] = 2.0**59
    rhoE[interior + (0,)] = rho
    got = oracle.mass_sd(rhoE, sd, ranks)
    assert got == _mass_restated(rho, nb, sd, ranks)
    assert oracle.mass_sd(rhoE, (1,) * dim, 1) == (oracle.mass2d(rhoE) if dim == 2 else oracle.mass3d(rhoE))


def test_mass_totals_differ_between_rank_counts(oracle):
    """The stiff state above gives different totals for one and two ranks, so the rank
    test is not passing on a sum that ignores the ranks."""
    nb = (17, 13)
    rhoE, _ = _state(2, nb, seed=1)
    rho = np.ones(nb)
    rho[0, 0] = 2.0**60
    rhoE[1:-1, 1:-1, 0] = rho
    one, two = oracle.mass_sd(rhoE, (2, 1), 1), oracle.mass_sd(rhoE, (2, 1), 2)
    assert one == 2.0**60 + 256 and two == 2.0**60  # 220 ones: rounded up once, or dropped in halves


@pytest.mark.parametrize("dim,nb,sd,ranks", [(2, (13, 11), (2, 3), 4), (2, (10, 10), (1, 3), 1),
                                             (3, (7, 6, 5), (2, 2, 2), 3)])
def test_csv_order_under_split_and_ranks(oracle, tmp_path, dim, nb, sd, ranks):
    rhoE, mom = _state(dim, nb, seed=3)
    path = tmp_path / "sd.csv"
    oracle.write_csv_sd(path, rhoE, mom, sd, ranks)
    lines = path.read_text().splitlines()
    assert lines[0] == ("x,y,rho,p,u,v" if dim == 2 else "x,y,z,rho,p,u,v,w")
    got = [tuple(int(x) for x in ln.split(",")[:dim]) for ln in lines[1:]]
    want = []
    for _, box in _order(nb, sd, ranks):
        (i0, ni), (j0, nj), (k0, nk) = box
        for i in range(i0, i0 + ni):
            for j in range(j0, j0 + nj):
                for k in range(k0, k0 + nk):
                    want.append((i, j) if dim == 2 else (i, j, k))
    assert got == want
    assert len(set(got)) == int(np.prod(nb))
    # the values of each row are the plain CSV's row of the same cell
    plain = tmp_path / "plain.csv"
    (oracle.write_csv2d if dim == 2 else oracle.write_csv3d)(plain, rhoE, mom)
    by_cell = {tuple(int(x) for x in ln.split(",")[:dim]): ln for ln in plain.read_text().splitlines()[1:]}
    assert all(by_cell[c] == ln for c, ln in zip(got, lines[1:]))
