"""Writes pocketsdr_amd/data/code_params.json: the ICD parameter tables of the
spreading codes used by the synthetic workloads, read as text from the
reference sources (run only where /root/reference is mounted).

  L5Q_XB_adv      XB code advance of GPS L5 Q5 (IS-GPS-705), PRN 1-32
  L6D_R_init      QZSS L6D shift-register initial states (IS-QZSS-L6), PRN 193-202
  L1CP_weil_idx   GPS L1C pilot Weil index (IS-GPS-800), PRN 1-32
  L1CP_ins_idx    GPS L1C pilot insertion index (IS-GPS-800), PRN 1-32
  E1B_hex         Galileo E1-B memory codes (Galileo OS SIS ICD), PRN 1-36,
                  4092 chips as 1023 hex digits

Source locations: src/sdr_code.c:173-215, 411-435; src/sdr_code_gal.c:18-.
"""
import json
import os
import re

REF = "/root/reference/src"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "pocketsdr_amd",
                   "data", "code_params.json")


def c_array(src, name):
    m = re.search(r"\b%s\[\]\s*=\s*\{(.*?)\};" % re.escape(name), src, re.S)
    body = re.sub(r"//[^\n]*", "", m.group(1))
    vals = []
    for tok in re.findall(r"\b(0x[0-9A-Fa-f]+|\d+)\b", body):
        if tok.startswith("0x"):
            vals.append(int(tok, 16))
        elif len(tok) > 1 and tok.startswith("0"):
            vals.append(int(tok, 8))          # C octal literal
        else:
            vals.append(int(tok))
    return vals


def main():
    code = open(os.path.join(REF, "sdr_code.c")).read()
    gal = open(os.path.join(REF, "sdr_code_gal.c")).read()
    m = re.search(r"code_gal_E1B\[\]\s*=\s*\{(.*?)\};", gal, re.S)
    hexs = "".join(re.findall(r'"([0-9A-Fa-f]*)"', m.group(1)))
    assert len(hexs) % 1023 == 0
    e1b = [hexs[i:i + 1023] for i in range(0, len(hexs), 1023)]
    out = dict(
        L5Q_XB_adv=c_array(code, "L5Q_XB_adv")[:32],
        L6D_R_init=c_array(code, "L6D_R_init")[:10],
        L1CP_weil_idx=c_array(code, "L1CP_weil_idx")[:32],
        L1CP_ins_idx=c_array(code, "L1CP_ins_idx")[:32],
        E1B_hex=e1b[:36],
    )
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", OUT, {k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
