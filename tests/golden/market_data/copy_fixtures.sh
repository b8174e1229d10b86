#!/bin/sh
# Copies two of the reference's small market-data series (market_data/ of the
# reference repository: raw float32 .bin, one-column _returns.csv and the full
# date,price,return,log_return .csv) here as loader fixtures.
set -e
SRC=/root/reference/market_data
DST=$(dirname "$0")
for s in spy_2020_covid spy_2008_crisis; do
    cp "$SRC/$s.bin" "$SRC/$s.csv" "$SRC/${s}_returns.csv" "$DST/"
done
