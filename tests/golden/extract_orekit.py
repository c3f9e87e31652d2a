"""Extract the Orekit degree/order-5 gravity vectors into a small fixture.

Inputs (read-only, present only in the development container):
  dust_transfer/crates/lightyear_odeint_rs/tests/data/orekit_dir_r6_5x5_v1.json
      5 body-fixed positions with Orekit 13.1.2 Holmes-Featherstone
      accelerations (hex binary64), the reference's gravity golden vectors
      (harness: LY/src/rhs.rs:892-997).
  dust_transfer/crates/two_phase_transfer_rs/data/spher_const/GO_CONS_GCF_2_DIR_R6_d15.txt
      ESA GOCE GO_CONS_GCF_2_DIR_R6 coefficients truncated to degree 15
      ("n m C S" fully normalised), licensed CC BY 4.0 (ESA).

Output: orekit_dir_r6_d5.json — positions, expected accelerations, model
constants and tolerances (hex strings kept bit-exact), plus the normalised
coefficients to degree 15 as data.  Run: python tests/golden/extract_orekit.py
"""
import json
import os

REF = "/root/reference/dust_transfer"
FIX = os.path.join(REF, "crates/lightyear_odeint_rs/tests/data/orekit_dir_r6_5x5_v1.json")
GFC = os.path.join(REF, "crates/two_phase_transfer_rs/data/spher_const/GO_CONS_GCF_2_DIR_R6_d15.txt")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "orekit_dir_r6_d5.json")


def main():
    with open(FIX) as fh:
        src = json.load(fh)
    coeffs = []
    with open(GFC) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) < 4:
                continue
            n, m = int(parts[0]), int(parts[1])
            c = float(parts[2].replace("D", "E").replace("d", "e"))
            s = float(parts[3].replace("D", "E").replace("d", "e"))
            coeffs.append([n, m, c, s])
    out = {
        "schema": "bqp_orekit_gravity_d5_v1",
        "source": "Orekit 13.1.2 Holmes-Featherstone comparator vectors shipped as test data "
                  "with the reference (orekit_dir_r6_5x5_v1.json); field ESA GOCE "
                  "GO_CONS_GCF_2_DIR_R6 (CC BY 4.0), fully normalised, tide free",
        "gm_m3_s2": src["model"]["gm_m3_s2"],
        "reference_radius_m": src["model"]["reference_radius_m"],
        "runtime_degree": 5,
        "absolute_tolerance_m_s2": src["evaluation"]["absolute_tolerance_m_s2"],
        "relative_tolerance": src["evaluation"]["relative_tolerance"],
        "cases": [
            {k: case[k] for k in ("name", "position_m", "orekit_noncentral_m_s2",
                                  "point_mass_m_s2", "full_m_s2")}
            for case in src["cases"]
        ],
        "coefficients_nmcs": coeffs,
    }
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
