"""Writes tests/golden/a513_cycles.json from the reference's A5.13 analysis (data/results/a513/analysis.json, JSON data
only): per point the first-accept run's counters per K ([flips, tests, anneals]), the cost per anneal of both solvers
(first-accept tests + flips; production Chord L + F), x_F, and the paired BER comparison of first-accept against the
production Chord in MDI (ratio, 95 % interval, BERs, uses). Run where the reference is mounted:
    python tests/golden/make_golden_cycles.py /root/reference"""
import json
import os
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
a = json.load(open(os.path.join(ref, 'data', 'results', 'a513', 'analysis.json')))
out = {}
for pt, r in a['points'].items():
    out[pt] = dict(method=r['method'], counters_fx=r['counters_fx'], cost_per_anneal=r['cost_per_anneal'],
                   x_F=r['x_F_vs_prod'], x_F_per_K=r['x_F_vs_prod_per_K'],
                   fx_prod={k: r['fx_prod'][k] for k in ('ratio', 'lo', 'hi', 'a_ber', 'b_ber', 'uses', 'a_bit_err',
                                                         'b_bit_err')})
json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'a513_cycles.json'), 'w'), indent=1)
print(sorted(out))
