#!/bin/sh
# Data fixtures: SNAP-format edge lists that the reference ships with its
# benchmark (benchmark/data/), copied unchanged so that GPU tests and the
# published-experiment bench workload can run on the real graphs where the
# reference tree is absent.
set -e
cd "$(dirname "$0")"
for f in bio-dmela ca-Erdos992 ca-GrQc ca-netscience in-arenas inf-power inf-euroroad; do
    cp /root/reference/benchmark/data/$f.txt $f.txt
done
