#!/bin/sh
# Data fixtures: SNAP-format edge lists that the reference ships with its
# benchmark (benchmark/data/), copied unchanged so that GPU tests can run the
# experiment harness on a real graph where the reference tree is absent.
set -e
cd "$(dirname "$0")"
cp /root/reference/benchmark/data/ca-netscience.txt ca-netscience.txt
cp /root/reference/benchmark/data/in-arenas.txt in-arenas.txt
