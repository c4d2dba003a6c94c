#!/bin/bash
# Mutation check of the oracle (test infrastructure): plants one plausible
# misreading of the reference at a time in a scratch copy of oracle/, builds
# it, and runs the oracle's CPU known-answer tests against it. Every mutant
# must fail at least one test; the script exits non-zero if one survives.
#   oracle/check_mutants.sh            (from the repository root, ~3 min)
set -u
ROOT=$(cd "$(dirname "$0")/.." && pwd)
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
# name | file | sed expression (first match only)
MUTANTS=(
  "view-prior-swapped|pm_oracle.cpp|0,/T(0.9) : T(0.1)/s//T(0.1) : T(0.9)/"
  "tau-0.6|pm_oracle.cpp|0,/T(0.8) \* std::exp/s//T(0.6) * std::exp/"
  "argmin-first-on-ties|pm_oracle.cpp|0,/if (c\[i\] <= m)/s//if (c[i] < m)/"
  "prior-from-far-flags|pm_oracle.cpp|0,/if (flag\[2 \* i\])/s//if (flag[2 * i + 1])/"
  "10-view-draws|pm_oracle.cpp|0,/for (int d = 0; d < 15; ++d) {/s//for (int d = 0; d < 10; ++d) {/"
  "fusion-average-by-count|fusion_oracle.cpp|0,/T(num) + T(1)/s//T(num)/"
  "fusion-reprojection-inclusive|fusion_oracle.cpp|0,/err < T(reproj_px)/s//err <= T(reproj_px)/"
  "fusion-count-exclusive|fusion_oracle.cpp|0,/if (num < geom_consistent)/s//if (num <= geom_consistent)/"
)
TESTS="tests/test_cpu_oracle_steps.py tests/test_cpu_fusion_twin.py tests/test_cpu_oracle.py"
survivors=0
for m in "${MUTANTS[@]}"; do
  IFS='|' read -r name file expr <<< "$m"
  d="$TMP/$name"
  mkdir -p "$d"
  cp -r "$ROOT/oracle" "$d/oracle"
  cp -r "$ROOT/tests" "$d/tests"
  ln -s "$ROOT/gmvs_amd" "$d/gmvs_amd"
  rm -f "$d"/oracle/*.so "$d"/oracle/*.o
  before=$(md5sum < "$d/oracle/$file")
  sed -i "$expr" "$d/oracle/$file"
  if [ "$before" == "$(md5sum < "$d/oracle/$file")" ]; then
    echo "$name: pattern not found in $file"; survivors=$((survivors + 1)); continue
  fi
  if ! (cd "$d/oracle" && make -s > /dev/null 2>&1); then
    echo "$name: build failed"; survivors=$((survivors + 1)); continue
  fi
  res=$(cd "$d" && python -m pytest -q -p no:cacheprovider $TESTS 2>&1 | tail -1)
  echo "$name: $res"
  case "$res" in *failed*) ;; *) survivors=$((survivors + 1)) ;; esac
done
echo "surviving mutants: $survivors"
exit $((survivors > 0))
