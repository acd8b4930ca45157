"""Writes exp_table.h: the 2^(k/128) table used by the device exp.

Entry k (0 <= k < 128) holds two 64-bit words:
  tail  = 2^(k/128) / H - 1 rounded to double, where H = 2^(k/128) rounded
          to double, so 2^(k/128) ~= H * (1 + tail) to about 106 bits;
  sbits = bits(H) - (k << 52) / 128, so that adding (k' << 45) for an
          integer k' = 128 m + k gives the bits of 2^m * H.
The values follow from the mathematics alone; they are computed here with
60-digit decimal arithmetic and rounded to double with Python's correctly
rounded float(str) conversion.

Run: python3 gen_exp_table.py > exp_table.h
"""
import struct
from decimal import Decimal, getcontext

getcontext().prec = 60


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def main():
    two = Decimal(2)
    lines = []
    for k in range(128):
        e = two ** (Decimal(k) / Decimal(128))
        h = float(str(e))
        tail = float(str(e / Decimal(h) - 1))
        sbits = (bits(h) - ((k << 52) // 128)) & 0xFFFFFFFFFFFFFFFF
        lines.append("    0x%016xull, 0x%016xull," % (bits(tail), sbits))
    print("// Generated by gen_exp_table.py: 2^(k/128) ~= asdouble(sbits + (k << 45)) * (1 + tail).")
    print("#pragma once")
    print("#include <cstdint>")
    print("namespace amcfd {")
    print("static constexpr uint64_t kExpTable[256] = {")
    print("\n".join(lines))
    print("};")
    print("}  // namespace amcfd")


if __name__ == "__main__":
    main()
