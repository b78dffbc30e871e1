"""Writes the committed fixtures under tests/golden/ from the reference's data files.

Run where the reference tree is present (it is read from /root/reference):
    python tests/golden/make_fixtures.py
It copies data, never code:
  * conf_04040408.npy      the 4x4x4x8 gauge field of conf_04040408.txt, in file
                           order (sites x 4 directions x 18 reals), float64;
  * rational/*             the nine parameter_rational.* coefficient files of
                           benchmark_hmc_dwf/ (plain text, copied verbatim);
  * known_answers.json     numbers printed by the reference's sample runs.
"""
import json
import os
import shutil

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    vals = np.loadtxt(os.path.join(REF, "conf_04040408.txt"), dtype=np.float64)
    assert vals.size == 4 * 4 * 4 * 8 * 4 * 18
    np.save(os.path.join(HERE, "conf_04040408.npy"), vals.reshape(-1, 4, 18))
    rdir = os.path.join(HERE, "rational")
    os.makedirs(rdir, exist_ok=True)
    src = os.path.join(REF, "benchmark_hmc_dwf")
    for f in sorted(os.listdir(src)):
        if f.startswith("parameter_rational."):
            shutil.copyfile(os.path.join(src, f), os.path.join(rdir, f))
    known = {
        "plaq_init_8x8x8x8_from_conf_04040408": {
            "value": 0.62512994697494162,
            "source": "sample_x86_corelib/run.log:725 (plaq = 6.2512994697494162e-01)",
        },
        "plaq_init_8x8x8x8_sample_1gpu": {
            "value": 0.62512994697494251,
            "source": "sample_1gpu/run.log:1558 (plaq = 6.2512994697494251e-01)",
        },
    }
    # energies printed by the 8^4 sample trajectory (sample_1gpu/run.log): the
    # reference's HMC_DWF_Nf2p1_eo.yaml on conf_04040408.txt tiled to 8^4,
    # RandomNumbers_parallel<Mseries> reset to 200000
    log = open(os.path.join(REF, "sample_1gpu", "run.log")).read().split("\n")
    want = [  # (key, line number, text before the number)
        ("H_kin_init", 1512, "H_kin        ="),
        ("H_Frational_init", 1523, "H_Frational  ="),
        ("H_Fratio_Nf2_init", 1529, "H_Fratio     ="),
        ("H_Fratio_Nf2prec_init", 1542, "H_Fratio     ="),
        ("H_Grectangle_init", 1554, "H_Grectangle ="),
        ("H_total_init", 1557, "H_total(init)  ="),
        ("H_kin_trial", 1957, "H_kin        ="),
        ("H_Frational_trial", 1968, "H_Frational  ="),
        ("H_Fratio_Nf2_trial", 1975, "H_Fratio     ="),
        ("H_Fratio_Nf2prec_trial", 1982, "H_Fratio     ="),
        ("H_Grectangle_trial", 1987, "H_Grectangle ="),
        ("H_total_trial", 1990, "H_total(trial) ="),
        ("plaq_trial", 1992, "plaq(trial)    ="),
        ("H_diff", 1994, "H(diff)       ="),
    ]
    for key, line, text in want:
        row = log[line - 1]
        assert text in row, (key, row)
        known[key] = {"value": float(row.split("=")[1]),
                      "source": "sample_1gpu/run.log:%d (%s)" % (line, row.strip())}
    # the same input file on 32x16x16x24 (4 ranks, grid [1, 1, 2, 2]; the parallel
    # generator's streams are per global (z, t), so one process draws the same numbers)
    name = "submit_miyabi/submit_4node.sh.o1137366"
    log = open(os.path.join(REF, name)).read().split("\n")
    want = [
        ("H_kin_init", 1515, "H_kin        ="),
        ("H_Frational_init", 1526, "H_Frational  ="),
        ("H_Fratio_Nf2_init", 1532, "H_Fratio     ="),
        ("H_Fratio_Nf2prec_init", 1545, "H_Fratio     ="),
        ("H_Grectangle_init", 1557, "H_Grectangle ="),
        ("H_total_init", 1560, "H_total(init)  ="),
        ("plaq_init", 1562, "plaq(init)     ="),
        ("H_kin_trial", 1960, "H_kin        ="),
        ("H_Frational_trial", 1971, "H_Frational  ="),
        ("H_Fratio_Nf2_trial", 1978, "H_Fratio     ="),
        ("H_Fratio_Nf2prec_trial", 1985, "H_Fratio     ="),
        ("H_Grectangle_trial", 1990, "H_Grectangle ="),
        ("H_total_trial", 1993, "H_total(trial) ="),
        ("plaq_trial", 1995, "plaq(trial)    ="),
        ("H_diff", 1997, "H(diff)       ="),
        ("trajectory_seconds", 2003, "each trajectory: total"),
    ]
    for key, line, text in want:
        row = log[line - 1]
        assert text in row, (key, row)
        v = row.split("total")[1].split("sec")[0] if key == "trajectory_seconds" else row.split("=")[1]
        known[key + "_32x16x16x24"] = {"value": float(v), "source": "%s:%d (%s)" % (name, line, row.strip())}
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(known, f, indent=1)


if __name__ == "__main__":
    main()
