"""Input files of the golden cases (shared by make_golden.py and the tests;
md_nucleation_amd.configs writes them deterministically)."""
import os

from md_nucleation_amd import configs


def write_case(name, d):
    m = configs.molecule_files(d)
    if name == "c1":
        # box 162 A: molecules inside the domain, in the 100-110 A blend shell
        # and outside it
        return configs.Condition(m["na"], m["methanol"], 30, "N2", 600, pressure=4e5,
                                 efield=(0.01, -0.02, 0.005)).write(os.path.join(d, "c1.cond"))
    if name == "pep":
        pep = configs.write_peptide(os.path.join(d, "pep40.txt"), seed=3, n=40)
        return configs.Condition(pep, m["water"], 20, "N2", 200, pressure=1e6, ion_potential="AMBER").write(
            os.path.join(d, "pep.cond"))
    if name in ("nvt_scale", "nvt_nh"):
        # the pep system under the relaxation thermostat of the ion
        pep = configs.write_peptide(os.path.join(d, "pep40.txt"), seed=3, n=40)
        nvt = "scale" if name == "nvt_scale" else "250"
        return configs.Condition(pep, m["water"], 20, "N2", 200, pressure=1e6, ion_potential="AMBER",
                                 nvt_ion=nvt).write(os.path.join(d, name + ".cond"))
    if name == "run":
        # MD::run on the pep system in a 94 A box (every molecule inside the
        # 110 A domain, so no rebuild draws random numbers): 100 relaxation
        # steps under velocity scaling, 300 NVE steps, Output every 50 steps,
        # Margin 0.5 A (several rebuilds, each logging the vapor).  No SHAKE H
        # line: MD::verlet runs no SHAKE, and the line would leave the X-H
        # bonds without any force
        pep = configs.write_peptide(os.path.join(d, "pep40.txt"), seed=3, n=40)
        return configs.Condition(pep, m["water"], 20, "N2", 200, pressure=1e6, ion_potential="AMBER",
                                 nvt_ion="scale", relax_steps=100, total_steps=300, observe=50,
                                 margin=0.5, shake_h=False).write(os.path.join(d, "run.cond"))
    if name == "bmh":
        nacl = configs.write_nacl_cluster(os.path.join(d, "nacl.txt"))
        return configs.Condition(nacl, m["water"], 1, "He", 1, ion_potential="Born-Mayer-Huggins-NaCl",
                                 interactions={}, shake_h=False).write(os.path.join(d, "bmh.cond"))
    if name in ("tersoff", "sw"):
        si = configs.write_si_cluster(os.path.join(d, "si.txt"))
        pot = "Tersoff" if name == "tersoff" else "Stilinger-Weber"
        return configs.Condition(si, m["water"], 1, "He", 1, ion_potential=pot, interactions={},
                                 shake_h=False).write(os.path.join(d, "si_%s.cond" % name))
    if name == "gg":
        # gas-gas LJ on: 200 N2 in a 94 A box (every molecule within 81 A of
        # the ion, so every blend weight is 1) around a 30-atom chain (the
        # reference's gg weight reads the ion's weights, GasGas.cpp:25, which
        # then hold at least the gas molecule's atom count)
        pep = configs.write_peptide(os.path.join(d, "pep30.txt"), seed=5, n=30)
        return configs.Condition(pep, m["methanol"], 5, "N2", 200, pressure=1e6, ion_potential="AMBER",
                                 interactions={"gg": "LJ", "gi": "LJ", "gv": "LJ", "vi": "LJcoul",
                                               "vv": "LJcoul"}).write(os.path.join(d, "gg.cond"))
    raise KeyError(name)
