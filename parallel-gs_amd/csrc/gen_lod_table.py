"""Writes lod_table.hpp: the smallest single-precision float >= 2^(k/128) for k = 0..128.

The mipmap LOD of a Q-derived level of detail is ceil(K - 2^(L+4) * log2|q|) in 1/16 units.
Writing |q| = m * 2^e with m in [1, 2), that integer is K - 2^(L+4) * e - floor(2^(L+4) * log2 m),
and floor(128 * log2 m) is the number of k in 1..127 with m >= 2^(k/128). Since 2^(k/128) is
irrational for 0 < k < 128, no float equals it, so m >= 2^(k/128) exactly when m is at least the
float above it: this table. The values are found with integer arithmetic only (M^128 against
2^(k + 23 * 128) for 24-bit mantissas M), so the table is exact.

Usage: python gen_lod_table.py > lod_table.hpp
"""
import struct


def threshold_mantissa(k: int) -> int:
    """Smallest M in [2^23, 2^24] with (M / 2^23)^128 >= 2^k."""
    target = 1 << (k + 23 * 128)
    lo, hi = 1 << 23, 1 << 24
    while lo < hi:
        mid = (lo + hi) // 2
        if mid ** 128 >= target:
            hi = mid
        else:
            lo = mid + 1
    return lo


def table():
    out = []
    for k in range(129):
        m = threshold_mantissa(k)
        out.append(m / float(1 << 23))  # exact: m has at most 25 bits and m / 2^23 <= 2
    return out


def main():
    vals = table()
    lines = [
        "// Generated by gen_lod_table.py: the smallest float >= 2^(k/128), k = 0..128 (exact, found",
        "// with integer arithmetic). floor(128 * log2 m) for m in [1, 2) is the largest k with",
        "// m >= kLodThreshold[k].",
        "#pragma once",
        "",
        "namespace pgs",
        "{",
        "constexpr float kLodThreshold[129] = {",
    ]
    for i in range(0, 129, 4):
        row = ", ".join(float(v).hex() + "f" for v in vals[i:i + 4])
        lines.append("\t" + row + ",")
    lines += ["};", "}", ""]
    print("\n".join(lines), end="")


if __name__ == "__main__":
    main()
