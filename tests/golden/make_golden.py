"""Extract the reference's own constant tables into tests/golden/reference_constants.json.

The reference stores no test vectors (every test draws from std::random_device);
its only fixed data are constant tables. This script reads them as text from the
reference tree and writes them as data (hex strings):

  include/host-arith/ntt-params.cuh:6-98     forward/inverse primitive 2^k-th roots
                                             and 2^-k, k = 0..28, Montgomery-form hex
  src/host-arith/curve/alt_bn128/alt_bn128_init.cu:40-61    moduli, R^2, R^3, inv
  alt_bn128_init.cu:116-132, :191-195                       curve b, twist, generators
  alt_bn128_init.cu:121-124                                 twist Frobenius constants
  alt_bn128_init.cu:250-254                                 pairing constants
  include/device-arith/curve/alt_bn128_t.cuh:11-45          device limb constants
  test/compile-check.cu:5-6                                 the two printed bigints

Run in the build container only (the reference tree is not present on GPU
machines): python tests/golden/make_golden.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def read(rel):
    with open(os.path.join(REF, rel)) as f:
        return f.read()


def table(src, name):
    m = re.search(name + r"\[[^\]]*\]\[65\]\s*=\s*\{(.*?)\};", src, re.S)
    return re.findall(r'"([0-9a-f]{64})"', m.group(1))


def main():
    ntt = read("include/host-arith/ntt-params.cuh")
    init = read("src/host-arith/curve/alt_bn128/alt_bn128_init.cu")
    dev = read("include/device-arith/curve/alt_bn128_t.cuh")
    cc = read("test/compile-check.cu")

    def bigint(label):
        m = re.search(re.escape(label) + r'\s*=\s*bigint[^(]*\("([0-9a-f]+)"\)', init)
        return m.group(1)

    def fq(label):
        m = re.search(re.escape(label) + r'\s*=\s*alt_bn128_Fq\("([0-9a-f]+)"\)', init)
        return m.group(1)

    out = {
        "_source": "constant tables of the reference tree, extracted as data by tests/golden/make_golden.py",
        "ntt_forward_root_of_unity_mont": table(ntt, "alt_bn128_forward_root_of_unity"),
        "ntt_inverse_root_of_unity_mont": table(ntt, "alt_bn128_inverse_root_of_unity"),
        "ntt_domain_scale_inverse_mont": table(ntt, "alt_bn128_domain_scale_inverse"),
        "modulus_r": bigint("alt_bn128_modulus_r"),
        "modulus_q": bigint("alt_bn128_modulus_q"),
        "fr_Rsquared": bigint("alt_bn128_Fr::Rsquared"),
        "fr_Rcubed": bigint("alt_bn128_Fr::Rcubed"),
        "fq_Rsquared": bigint("alt_bn128_Fq::Rsquared"),
        "fq_Rcubed": bigint("alt_bn128_Fq::Rcubed"),
        "fr_inv": re.search(r"alt_bn128_Fr::inv\s*=\s*0x([0-9a-f]+)", init).group(1),
        "fq_inv": re.search(r"alt_bn128_Fq::inv\s*=\s*0x([0-9a-f]+)", init).group(1),
        "fr_root_of_unity": re.search(r'alt_bn128_Fr::root_of_unity\s*=\s*alt_bn128_Fr\("([0-9a-f]+)"\)', init).group(1),
        "fr_multiplicative_generator": re.search(r'alt_bn128_Fr::multiplicative_generator\s*=\s*alt_bn128_Fr\("([0-9a-f]+)"\)', init).group(1),
        "coeff_b": fq("alt_bn128_coeff_b"),
        "twist_mul_by_q_X": re.findall(r'alt_bn128_twist_mul_by_q_X\s*=\s*alt_bn128_Fq2\(alt_bn128_Fq\("([0-9a-f]+)"\),\s*alt_bn128_Fq\("([0-9a-f]+)"\)', init)[0],
        "twist_mul_by_q_Y": re.findall(r'alt_bn128_twist_mul_by_q_Y\s*=\s*alt_bn128_Fq2\(alt_bn128_Fq\("([0-9a-f]+)"\),\s*alt_bn128_Fq\("([0-9a-f]+)"\)', init)[0],
        "g2_one": re.findall(r'G2_one\s*=\s*alt_bn128_G2\(alt_bn128_Fq2\(alt_bn128_Fq\("([0-9a-f]+)"\),\s*alt_bn128_Fq\("([0-9a-f]+)"\)\),\s*'
                             r'alt_bn128_Fq2\(alt_bn128_Fq\("([0-9a-f]+)"\),\s*alt_bn128_Fq\("([0-9a-f]+)"\)\)', init)[0],
        "ate_loop_count": bigint("alt_bn128_ate_loop_count"),
        "final_exponent": re.search(r'alt_bn128_final_exponent\s*=\s*bigint<[^>]*>\("([0-9a-f]+)"\)', init).group(1),
        "final_exponent_z": bigint("alt_bn128_final_exponent_z"),
        "device_M0_fq": re.search(r"ALT_BN128_M0\s*=\s*0x([0-9a-f]+)", dev).group(1),
        "device_m0_fr": re.search(r"ALT_BN128_m0\s*=\s*0x([0-9a-f]+)", dev).group(1),
        "compile_check_bigints": re.findall(r'"([0-9a-fA-F]{16,})"', cc),
    }
    for k in ("ntt_forward_root_of_unity_mont", "ntt_inverse_root_of_unity_mont", "ntt_domain_scale_inverse_mont"):
        assert len(out[k]) == 29, (k, len(out[k]))
    with open(os.path.join(HERE, "reference_constants.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "reference_constants.json"))


if __name__ == "__main__":
    main()
