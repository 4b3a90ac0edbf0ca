"""Extract the pow_of_gsq table (RandomNumbers/RNG_class.f90:61-123) as a data fixture.

Run once in a container that has the reference tree; the JSON it writes is committed."""
import json
import os
import re
import sys

SRC = os.environ.get("SCONE_REF", "/root/reference") + "/RandomNumbers/RNG_class.f90"


def main():
    text = open(SRC).read()
    block = text[text.index("pow_of_gsq = ["):]
    block = block[: block.index("]") + 1]
    vals = [int(v) for v in re.findall(r"(\d+)_int64", block)]
    assert len(vals) == 63, len(vals)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_pow_of_gsq.json")
    json.dump({"source": "RandomNumbers/RNG_class.f90:61-123", "g": 2806196910506780709,
               "pow_of_gsq": vals}, open(out, "w"), indent=1)
    print(out)


if __name__ == "__main__":
    sys.exit(main())
