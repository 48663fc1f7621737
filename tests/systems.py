"""Test molecules: the geometries of NOPT's example inputs (Angstrom) and the
benchmark systems."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# examples/c2h4_2_2/cas.inp:9-14
C2H4 = (["C", "C", "H", "H", "H", "H"],
        [[0, 0, -0.6695], [0, 0, 0.6695],
         [0, -0.928925576685832, 1.23207734844099], [0, 0.928925576685832, -1.23207734844099],
         [0, -0.928925576685832, -1.23207734844099], [0, 0.928925576685832, 1.23207734844099]])
# examples/cis/c6h8_cis.inp:9-22
C6H8 = (["C"] * 6 + ["H"] * 8,
        [[0.000000e+00, 0.000000e+00, 0.0], [1.348000e+00, 0.000000e+00, 0.0],
         [2.082000e+00, 1.271325e+00, 0.0], [3.430000e+00, 1.271325e+00, 0.0],
         [4.164000e+00, 2.542651e+00, 0.0], [5.512000e+00, 2.542651e+00, 0.0],
         [-5.535000e-01, -9.586901e-01, 0.0], [-5.535000e-01, 9.586901e-01, 0.0],
         [1.901500e+00, -9.586901e-01, 0.0], [1.528500e+00, 2.230015e+00, 0.0],
         [3.983500e+00, 3.126352e-01, 0.0], [3.610500e+00, 3.501341e+00, 0.0],
         [6.065500e+00, 1.583960e+00, 0.0], [6.065500e+00, 3.501341e+00, 0.0]])
# examples/spin_proj/guanine_ss4.inp:9-24
GUANINE = (["C", "N", "H", "C", "C", "N", "C", "O", "N", "H", "C", "N", "N", "H", "H", "H"],
           [[0.0, -2.827132420000, -5.424758309000], [0.0, -1.758758998000, -4.575570863000],
            [0.0, -0.822927268000, -4.830881393000], [0.0, -2.249041855000, -3.297846026000],
            [0.0, -3.643437847000, -3.417138842000], [0.0, -3.933900153000, -4.740985261000],
            [0.0, -4.423912346000, -2.237455129000], [0.0, -5.641774093000, -2.294076817000],
            [0.0, -3.786038908000, -1.046032887000], [0.0, -4.300096075000, -0.223447794000],
            [0.0, -2.423974018000, -0.994413550000], [0.0, -1.687653945000, -2.083169454000],
            [0.0, -1.803572326000, 0.229529273000], [0.0, -2.763278662000, -6.502824696000],
            [0.0, -2.333561673000, 1.041945018000], [0.0, -0.835058139000, 0.282350383000]])
# examples/dmrg/pBQ/pBQ_2A_RHF.inp:8-19 (cc-pVDZ / cc-pVDZ-JKFIT, D5=1); NOPT's
# RHF orbitals of this input are examples/dmrg/pBQ/pBQ_2B2_RHF.orb, kept as
# tests/golden/pBQ_2B2_RHF.orb
PBQ = (["C", "C", "O", "O", "C", "C", "C", "C", "H", "H", "H", "H"],
       [[-1.4365365158, 0.0, 0.0], [1.4365365158, 0.0, 0.0], [-2.6866039788, 0.0, 0.0],
        [2.6866039788, 0.0, 0.0], [-0.6867858251, -1.2538739222, 0.1540165049],
        [0.6867858251, -1.2538739222, -0.1540165049], [-0.6867858251, 1.2538739222, -0.1540165049],
        [0.6867858251, 1.2538739222, 0.1540165049], [-1.2757216894, -2.1643648836, 0.1703096948],
        [1.2757216894, -2.1643648836, -0.1703096948], [-1.2757216894, 2.1643648836, -0.1703096948],
        [1.2757216894, 2.1643648836, 0.1703096948]])
PBQ_ORB = os.path.join(HERE, "golden", "pBQ_2B2_RHF.orb")
# examples/dmrg/4ene/4ene_RHF.inp:8-27 (octatetraene, cc-pVDZ / cc-pVDZ-JKFIT,
# D5=1); NOPT's RHF orbitals of this input are examples/dmrg/4ene/4ene_RHF.orb,
# kept as tests/golden/4ene_RHF.orb
ENE4 = (["C"] * 8 + ["H"] * 10,
        [[4.33581459, -0.21013946, 0.0], [-4.33581459, 0.21013946, 0.0],
         [3.13707423, 0.41982979, 0.0], [-3.13707423, -0.41982979, 0.0],
         [1.84790846, -0.28094470, 0.0], [-1.84790846, 0.28094470, 0.0],
         [0.64299768, 0.34910409, 0.0], [-0.64299768, -0.34910409, 0.0],
         [4.40307799, -1.30401673, 0.0], [-4.40307799, 1.30401673, 0.0],
         [5.27492320, 0.35051356, 0.0], [-5.27492320, -0.35051356, 0.0],
         [3.10371272, 1.51717565, 0.0], [-3.10371272, -1.51717565, 0.0],
         [1.87744497, -1.37938629, 0.0], [-1.87744497, 1.37938629, 0.0],
         [0.61477995, 1.44763979, 0.0], [-0.61477995, -1.44763979, 0.0]])
ENE4_ORB = os.path.join(HERE, "golden", "4ene_RHF.orb")

# RI-RHF energies printed by NOPT (cc-pVDZ / cc-pVDZ-RI, D5=1)
GOLDEN_E = {
    "C2H4": -78.0411981212,    # examples/c2h4_2_2/benchmark.log:398
    "C6H8": -231.8259634059,   # examples/cis/benchmark.log:757
    "guanine": -539.4212820218,  # examples/spin_proj/guanine_ss4.log:1044
}
# NOPT's state-averaged CASSCF(2,2) of C2H4 (examples/c2h4_2_2/cas.inp: two
# singlets, weights 1 1, cc-pVDZ / cc-pVDZ-RI, D5=1): the CI roots at the RHF
# orbitals (benchmark.log:468, 473), the averaged energy (benchmark.log:507)
# and the states at convergence (benchmark.log:542-543). NOPT stops at an
# orbital gradient of 1e-5, so only the average is fixed to the printed
# digits; each state moves linearly with the orbital error.
CAS_C2H4 = {"ci_rhf": [-78.0595141689, -77.6877369591], "E_avg": -77.8755259203,
            "states": [-78.0575925136, -77.6934593271]}

# NOPT's state-averaged CASSCF(8,8) of octatetraene (examples/dmrg/4ene/
# 4ene_CAS.inp: three singlets, weights 1 1 1, from 4ene_RHF.orb): the CI
# roots at the file's orbitals (4ene_CAS.log:1326-1328), the averaged energy
# (4ene_CAS.log:1349, orbital gradient 7.8e-7) and the states at convergence
# (4ene_CAS.log:1408-1410)
CAS_4ENE = {"ci_file": [-308.7964933374, -308.6103724502, -308.5702725368], "E_avg": -308.6944759121,
            "states": [-308.8261097128, -308.6493233879, -308.6079946356]}

# NOPT's CASSCF of the p-benzoquinone anion (examples/dmrg/pBQ/pBQ_2A_CAS.inp:
# charge -1, doublets, two states, from pBQ_2B2_RHF.orb): the CI roots at the
# file's orbitals (pBQ_2A_CAS.log:1065-1066)
CAS_PBQ = {"ci_file": [-379.2978536692, -379.2148560831]}

SYSTEMS = {"C2H4": C2H4, "C6H8": C6H8, "guanine": GUANINE, "pBQ": PBQ}
