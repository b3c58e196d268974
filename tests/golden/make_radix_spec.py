"""Writes radix_partition_spec.json: the cases of operations/RadixPartition_test.cpp:106-168.

Each case lists the input column and, per partition, the expected (value, tuple_id) multiset,
exactly as the reference test builds them (expected_partitions[i % partitions].emplace(i, i),
expected_partitions[0].emplace(0, i), and the 0/1 interleave).  Sizes are capped at 1000 here
to keep the fixture small; the tests also run the 100000-row sizes from the rule directly.
"""
import json
import os

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "radix_partition_spec.json")


def main():
    cases = []
    for bits in (1, 3, 7, 12):
        parts = 1 << bits
        for size in (10, 20, 40, 100, 1000):
            simple = {"name": "SimpleTests", "radix_bits": bits, "values": list(range(size)),
                      "partitions": {}}
            for i in range(size):
                simple["partitions"].setdefault(str(i % parts), []).append([i, i])
            cases.append(simple)
            zero = {"name": "AllZero", "radix_bits": bits, "values": [0] * size,
                    "partitions": {"0": [[0, i] for i in range(size)]}}
            cases.append(zero)
            inter = {"name": "ZeroInterleaveOne", "radix_bits": bits, "values": [], "partitions": {"0": [], "1": []}}
            for i in range(size // 2):
                inter["values"] += [0, 1]
                inter["partitions"]["0"].append([0, 2 * i])
                inter["partitions"]["1"].append([1, 2 * i + 1])
            cases.append(inter)
    with open(OUT, "w") as f:
        json.dump(cases, f)


if __name__ == "__main__":
    main()
