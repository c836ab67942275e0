"""Writes log2_table.inc: entry i = floor(log2(1 + i/256) * 2**16), computed
exactly with integers: floor(2**16 * log2(m)) = bitlen(m**(2**16)) - 1."""
import os

def table():
    out = []
    for i in range(256):
        m = 256 + i
        out.append((m ** 65536).bit_length() - 1 - 8 * 65536)
    return out

if __name__ == "__main__":
    t = table()
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "log2_table.inc")
    with open(path, "w") as f:
        f.write("/* generated by gen_log2_table.py: floor(log2(1 + i/256) * 65536) */\n")
        for k in range(0, 256, 12):
            f.write("    " + ", ".join(str(v) for v in t[k:k + 12]) + ",\n")
