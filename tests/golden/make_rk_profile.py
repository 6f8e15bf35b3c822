"""Writes rk_wellbore_profile.json from the reference's verification data
(mathematica_notebooks/RK_Verification/): the Runge-Kutta elastoplastic
wellbore profile (ElastoPlasticRKdata.txt, every 25th of 10,000 rows) and the
reference IntPointsFEM CPU result sampled along y = 0
(ElastoPlasticIntPoints-CPU.csv, every 5th of 1,001 rows).  Case
(main.cpp:374-445): E = 2000, nu = 0.2, c = 5, phi = psi = 20 deg, wall stress
increment -10, outer -0.00625.

    python tests/golden/make_rk_profile.py   (reads /root/reference)
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/mathematica_notebooks/RK_Verification"

if __name__ == "__main__":
    full = np.loadtxt(os.path.join(SRC, "ElastoPlasticRKdata.txt"), skiprows=1)[::-1]
    plastic = (np.abs(full[:, 8]) + np.abs(full[:, 9])) > 1e-9
    rk = full[::25]
    ip = np.genfromtxt(os.path.join(SRC, "ElastoPlasticIntPoints-CPU.csv"), delimiter=",", names=True)[::5]
    out = {
        "case": {"E": 2000.0, "nu": 0.2, "cohesion": 5.0, "phi_deg": 20.0, "psi_deg": 20.0,
                 "sigma_n_inner": -10.0, "sigma_n_outer": -0.00625, "r_in": 0.1, "r_out": 4.0,
                 "rk_plastic_radius": float(full[plastic, 0].max())},
        "rk": {k: rk[:, i].round(9).tolist() for i, k in enumerate(
            ["radius", "u", "sigma_rr", "sigma_tt", "sigma_zz", "eps_rr", "eps_tt", "eps_zz", "eps_p_rr",
             "eps_p_tt", "eps_p_zz"])},
        "intpoints_cpu": {"x": ip["Points0"].tolist(), "sigma_xx": ip["Stress0"].tolist(),
                          "sigma_yy": ip["Stress4"].tolist(), "sigma_zz": ip["Stress8"].tolist(),
                          "u_x": ip["Displacement0"].tolist(), "eps_p_xx": ip["StrainPlastic0"].tolist(),
                          "eps_p_yy": ip["StrainPlastic4"].tolist()},
    }
    with open(os.path.join(HERE, "rk_wellbore_profile.json"), "w") as f:
        json.dump(out, f)
    print(len(out["rk"]["radius"]), "RK rows,", len(out["intpoints_cpu"]["x"]), "IntPoints rows")
