"""Copy the reference's numeric test fixtures (data files, not code) into tests/golden/reference_data.

Run once in the development container where the reference tree is mounted:
    python tests/golden/import_reference_fixtures.py /root/reference
The files are the known-answer data that lao19881213/purify-1's own Catch2 tests read
(cpp/tests/kernels.cc, operators.cc, measurement_operator.cc, algo_factory.cc).  They are GPL-2 data from that
project; PROVENANCE.md in the output folder records where each came from.  The numeric vectors
that the reference's wproj_utils test (cpp/tests/wproj_utils.cc:217-289) holds inline are written
to convolution_known_answers.json.
"""
import json
import os
import shutil
import sys

FILES = {
    "kernels": ["kernel_x_data", "kernel_kb_data", "kernel_gauss_data", "kernel_pswf_data",
                "correction_x_data", "correction_kb_data", "correction_gauss_data", "correction_pswf_data"],
    "operators": ["u_data", "v_data", "direct_input_data", "expected_direct_data", "indirect_input_data",
                  "expected_indirect_data", "expected_S_data"],
    "measurement_operator": ["u", "v", "kb_exact_data", "pswf_exact_data", "gauss_exact_data"],
    # the PADMM factory fixture (cpp/tests/algo_factory.cc:20-67): input visibilities, SOPT's solution
    # and residual images and the run's parameters
    "padmm": ["input_data.vis", "solution.fits", "residual.fits", "test_parameters"],
}

# Inline data of cpp/tests/wproj_utils.cc:236-244 (4x4) and 271-282 (5x5): G row, chirp row and
# the expected row-wise convolution, 16 and 25 real values each.
CONV = {
    "even_4x4": {
        "G": [0.7634463, 0.8537454, 0.89558853, 0.71197476, 0.07342149, 0.19600616, 0.21977598, 0.3802055,
              0.07577487, 0.94388022, 0.5813741, 0.51562299, 0.433173, 0.34491308, 0.55125663, 0.9279308],
        "C": [0.13239976, 0.01737199, 0.44752763, 0.33047119, 0.71925933, 0.40904651, 0.80406293,
              0.95091037, 0.61481599, 0.39252926, 0.22208961, 0.36976165, 0.65017947, 0.22863747,
              0.87700107, 0.06965994],
        "K": [1.56301893, 0.69579421, 2.55646966, 2.80081109, 1.73328232, 0.95357603, 2.12264685,
              2.45700481, 1.91327398, 1.27415219, 2.06766882, 2.84902461, 2.33590802, 1.64204955,
              3.13634125, 3.72991149],
        "N": 4,
    },
    "odd_5x5": {
        "G": [0.84416411, 0.95108536, 0.66690087, 0.92687553, 0.82640953, 0.05214219, 0.23629002,
              0.0456125, 0.31058604, 0.20523312, 0.33501332, 0.2103377, 0.67046083, 0.68790078,
              0.84890013, 0.62349394, 0.18208434, 0.91033138, 0.34436924, 0.11369944, 0.31609394,
              0.45869556, 0.63513753, 0.73314836, 0.51449701],
        "C": [0.30409785, 0.20084303, 0.32436196, 0.49613954, 0.29283696, 0.54984908, 0.60762237,
              0.74812515, 0.96318413, 0.80115342, 0.85182094, 0.33545398, 0.1658094, 0.31987092,
              0.41907388, 0.22498119, 0.60000475, 0.67336284, 0.45901998, 0.550584, 0.90333773,
              0.80207468, 0.20964248, 0.84893657, 0.6063695],
        "K": [6.2710243, 4.53980133, 3.21624657, 3.82809291, 4.6633187, 6.09692671, 4.24215225,
              3.24627495, 4.52701773, 5.53094158, 3.27058866, 2.46446407, 1.58955594, 2.41529522,
              2.84256957, 4.63801315, 3.28906792, 2.81268103, 2.74098252, 3.65394543, 5.83657047,
              4.21082195, 3.19483206, 3.67326191, 4.7494392],
        "N": 5,
    },
}


def main(ref_root):
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_data")
    os.makedirs(out, exist_ok=True)
    lines = ["# Provenance of the reference fixtures", "",
             "Copied verbatim from lao19881213/purify-1 (PURIFY 2.0.0, GPL-2) by "
             "`tests/golden/import_reference_fixtures.py`.", ""]
    for sub, names in FILES.items():
        os.makedirs(os.path.join(out, sub), exist_ok=True)
        for n in names:
            src = os.path.join(ref_root, "data", "expected", sub, n)
            shutil.copyfile(src, os.path.join(out, sub, n))
            lines.append(f"- `{sub}/{n}` <- `data/expected/{sub}/{n}`")
    os.makedirs(os.path.join(out, "images"), exist_ok=True)
    shutil.copyfile(os.path.join(ref_root, "data", "images", "M31.fits"), os.path.join(out, "images", "M31.fits"))
    lines.append("- `images/M31.fits` <- `data/images/M31.fits`")
    with open(os.path.join(out, "convolution_known_answers.json"), "w") as f:
        json.dump(CONV, f, indent=1)
    lines.append("- `convolution_known_answers.json` <- numeric vectors inline in "
                 "`cpp/tests/wproj_utils.cc:236-244,271-282`")
    lines += ["", "Readers: `tests/golden/fixtures.py`.  The reference tests that consume them are "
              "`cpp/tests/kernels.cc:35-73`, `cpp/tests/operators.cc:65-101`, "
              "`cpp/tests/measurement_operator.cc:14-84`, `cpp/tests/algo_factory.cc:20-67` and "
              "`cpp/tests/wproj_utils.cc:217-289`."]
    with open(os.path.join(out, "PROVENANCE.md"), "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
