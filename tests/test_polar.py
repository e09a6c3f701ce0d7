"""Polar CSV output (post::toCsv, Polar.cpp:126-163) against the polar the
reference wrote itself (tests/golden/naca0012_re1e6_sst_polar.csv): the rows
parsed back and formatted again reproduce the file byte for byte. CPU only."""
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "naca0012_re1e6_sst_polar.csv")
STATUS = {"converged": 0, "iteration-limit": 1, "diverged": 2}


def parse(text):
    header, rows = {}, []
    for line in text.splitlines():
        if line.startswith("# ") and "," in line:
            k, v = line[2:].split(",", 1)
            header[k] = v
        elif line and not line.startswith("#") and not line.startswith("alpha_deg"):
            rows.append(line.split(","))
    return header, rows


def test_polar_csv_reproduces_the_reference_file(G):
    text = open(GOLDEN).read()
    header, rows = parse(text)
    pts = []
    for r in rows:
        p = G.PolarPoint()
        p.alpha_deg, p.cl, p.cd, p.cd_pressure, p.cd_friction, p.cm, p.l_over_d = map(float, r[:7])
        p.upper_separation = float(r[7]) if r[7] else -1.0
        p.lower_separation = float(r[8]) if r[8] else -1.0
        p.status = STATUS[r[9]]
        p.iterations = int(r[10])
        p.continuity = float(r[11])
        pts.append(p)
    out = G.polar_csv(pts, section=header["section"], mesh=header["mesh"], reynolds=float(header["reynolds"]),
                      speed=float(header["freestream_speed_mps"]), chord=float(header["chord_m"]),
                      moment_reference=float(header["moment_reference_xc"]),
                      continued=header["continued_between_points"] == "yes")
    assert out == text


def test_polar_csv_formats_status_and_attached_separation(G):
    p = G.PolarPoint(alpha_deg=-2.5, cl=-0.25, cd=0.0125, cd_pressure=0.005, cd_friction=0.0075, cm=0.001,
                     l_over_d=-20.0, upper_separation=-1.0, lower_separation=0.5, status=2, iterations=17,
                     continuity=float("inf"))
    out = G.polar_csv([p], section="NACA 2412", mesh="Fine", reynolds=2.5e6, speed=50.0, chord=1.0,
                      moment_reference=0.25, continued=False).splitlines()
    assert out[1] == "# section,NACA 2412" and out[3] == "# reynolds,2.5e+06"
    assert out[7] == "# continued_between_points,no"
    assert out[-1] == "-2.500000,-0.25000000,0.01250000,0.00500000,0.00750000,0.00100000,-20.000000,,0.500000,diverged,17,inf"
