#!/usr/bin/env python3
"""Extract the reference's published verification results into a JSON
fixture (data only: the numbers the reference's own verification runs
produced, each with the file and line it comes from).

Run in the authoring container, where the reference is mounted:
    python tests/golden/make_reference_verification.py
"""
import json
import os

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_verification.json")


def row(path, line):
    with open(os.path.join(REF, path)) as fh:
        lines = fh.read().splitlines()
    return [float(x) for x in lines[line - 1].split()]


def main():
    data = {
        "stokes_drag_tau0p7": {
            "source": "verification/3D_stokesDragForce-tau0p7/ForceError.dat:2",
            "deck": "verification/3D_stokesDragForce-tau0p7/in.lbmpsm",
            "tau": row("verification/3D_stokesDragForce-tau0p7/ForceError.dat", 2)[0],
            "force_error_percent": row("verification/3D_stokesDragForce-tau0p7/ForceError.dat", 2)[1],
        },
        "torque_shear_tau0p7": {
            "source": "verification/3D_torque-tau0p7/ErrorSummary.dat:2",
            "deck": "verification/3D_torque-tau0p7/in.lbmpsm",
            "stresslet_error_percent": row("verification/3D_torque-tau0p7/ErrorSummary.dat", 2)[1],
            "torque_error_percent": row("verification/3D_torque-tau0p7/ErrorSummary.dat", 2)[2],
        },
        "stresslet_free_rotation_tau0p7": {
            "source": "verification/3D_stresslet-tau0p7/ErrorSummary.dat:2",
            "stresslet_error_percent": row("verification/3D_stresslet-tau0p7/ErrorSummary.dat", 2)[1],
            "omega_error_percent": row("verification/3D_stresslet-tau0p7/ErrorSummary.dat", 2)[2],
        },
        "drag_coefficient_2d": {
            "source": "verification/2D_dragForce/dragCoefficients.dat:2-4",
            "Re": [row("verification/2D_dragForce/dragCoefficients.dat", i)[0] for i in (2, 3, 4)],
            "Cd": [row("verification/2D_dragForce/dragCoefficients.dat", i)[1] for i in (2, 3, 4)],
        },
        "suspension_viscosity": {
            "source": "verification/shearedSuspensions/phi-viscosity.dat:2-6",
            "phi": [row("verification/shearedSuspensions/phi-viscosity.dat", i)[0] for i in range(2, 7)],
            "eta_rel": [row("verification/shearedSuspensions/phi-viscosity.dat", i)[1] for i in range(2, 7)],
        },
        "benchmark_wall_time_s": {
            "source": "benchmark/benchmark-results.dat:2-7",
            "rows": [row("benchmark/benchmark-results.dat", i) for i in range(2, 8)],
        },
    }
    with open(OUT, "w") as fh:
        json.dump(data, fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
