#!/bin/sh
# Per-trajectory batch loop of TD_GPU_code/run.sh:3-19 without the LSF
# templating: every Trajectories/<name>.traj is run through teuk_run (all
# m-modes of the parameter file), writing hm0<m>_<name>.dat / hmm<|m|>_<name>.dat
# into the current directory.
#
#   run.sh [params.txt]          (run from the directory holding Trajectories/)
HERE=$(cd "$(dirname "$0")" && pwd)
BIN="$HERE/../bin/teuk_run"
PARAMS=${1:-}
for file in Trajectories/*.traj; do
    [ -e "$file" ] || { echo "no Trajectories/*.traj here" >&2; exit 1; }
    bfile=$(basename "$file" .traj)
    echo "$bfile"
    if [ -n "$PARAMS" ]; then
        "$BIN" -p "$PARAMS" -t Trajectories -o . "$bfile" || exit $?
    else
        "$BIN" -t Trajectories -o . "$bfile" || exit $?
    fi
done
