"""Host parts of the command-line driver: the reference's momentum-list format
(lib/catchOpts.cpp:106-200; a line such as those of tests/mom_pipi.dat) and the site phases."""
import numpy as np
import pytest

from contractions_code_scattering_amd import cli


def test_read_momenta(tmp_path):
    p = tmp_path / "mom.dat"
    p.write_text("1 0 0\t-1 0 0\t1 0 0\t-1 0 0\n0 1 0\t0 -1 0\t0 1 0\t0 -1 0\n")
    assert cli.read_momenta(str(p), 4, 2) == [[(1, 0, 0), (-1, 0, 0), (1, 0, 0), (-1, 0, 0)],
                                              [(0, 1, 0), (0, -1, 0), (0, 1, 0), (0, -1, 0)]]
    with pytest.raises(SystemExit):
        cli.read_momenta(str(p), 4, 3)  # wrong number of combinations
    with pytest.raises(SystemExit):
        cli.read_momenta(str(p), 2, 2)  # wrong number of columns


def test_site_phases():
    nx, ny, nz = 4, 3, 2
    ph = cli.site_phases([(1, 2, 1), (1, 2, 1)], [-1, 1], nx, ny, nz)
    assert ph.shape == (2, 3 * nx * ny * nz)
    # site x + nx (y + ny z), three colours each
    x, y, z = 3, 2, 1
    k = 3 * (x + nx * (y + ny * z))
    want = np.exp(-2j * np.pi * (x * 1 / nx + y * 2 / ny + z * 1 / nz))
    assert np.allclose(ph[0, k:k + 3], want)
    assert np.allclose(ph[1], np.conj(ph[0]))


def test_parse_needs_elementals():
    with pytest.raises(SystemExit):
        cli.parse(["rho", "--propname", "p"])
    # the derivative types need links; each diagram its driver's number of momenta
    with pytest.raises(SystemExit):
        cli.parse(["rho", "--propname", "p", "--eigename", "e", "--momname", "m"])
    with pytest.raises(SystemExit):
        cli.parse(["sigmapi", "--propname", "p", "--eigename", "e", "--momname", "m", "--linkname", "l"])
    a = cli.parse(["sigmapi", "--propname", "p", "--eigename", "e", "--momname", "m", "--linkname", "l",
                   "--Nmom", "2"])
    assert a.linkname == "l" and a.Nmom == 2
    assert cli.parse(["pipi_I1", "--propname", "p", "--eigename", "e", "--momname", "m", "--Nmom", "4"]).Nmom == 4
