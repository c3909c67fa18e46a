#!/bin/bash
# CPU test suite under AddressSanitizer + UndefinedBehaviorSanitizer: the library's host code (builder, collections,
# C ABI, autotuner) and the oracle are rebuilt with both sanitizers and loaded in place of the regular builds;
# Python itself is not instrumented, so the sanitizer runtimes are preloaded. Leak checking is off (the
# interpreter's own allocations would be reported). Usage: bash tests/run_sanitized.sh [extra pytest args]
set -euo pipefail
ROOT="$(cd "$(dirname "${BASH_SOURCE[0]}")/.." && pwd)"
cd "$ROOT"
python -c "from blazert_amd import build as B; B.build_sanitized(verbose=True)"
make -s -C oracle sanitize
export BLAZERT_AMD_LIB="$ROOT/blazert_amd/_asan/libblazert_amd.so"
export BLAZERT_ORACLE_LIB="$ROOT/oracle/_asan/liboracle.so"
export BLAZERT_ORACLE_F32_LIB="$ROOT/oracle/_asan/liboracle_f32.so"
# reports go to files (pytest captures the test process's stderr), printed below
LOGS="$(mktemp -d /tmp/brt_sanitize.XXXXXX)"
export ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:log_path=$LOGS/asan"
export UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:log_path=$LOGS/ubsan"
rc=0
LD_PRELOAD="$(gcc -print-file-name=libasan.so):$(gcc -print-file-name=libubsan.so)" \
  python -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@" || rc=$?
if compgen -G "$LOGS/*" > /dev/null; then
  cat "$LOGS"/*
  echo "sanitizer reports above"
  rc=1
fi
rm -rf "$LOGS"
exit $rc
