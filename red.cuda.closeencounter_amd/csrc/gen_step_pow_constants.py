"""Prints the double-double constants of step_pow.hpp: 1/(2k+1) for the atanh series
(k = 0..24) and 1/i! for the exponential's Taylor sum (i = 0..15), each as hi + lo with
hi = the double nearest the value and lo = the double nearest the rest, from Python's
decimal module at 60 digits."""
from decimal import Decimal, getcontext

getcontext().prec = 60


def dd(x: Decimal):
    hi = float(x)
    lo = float(x - Decimal(hi))
    return hi, lo


def emit(name, values):
    print(f"constexpr double {name}[{len(values)}][2] = {{")
    for v in values:
        hi, lo = dd(v)
        print(f"    {{{hi.hex()}, {lo.hex()}}},")
    print("};")


emit("INV_ODD", [Decimal(1) / Decimal(2 * k + 1) for k in range(25)])
fact = [Decimal(1)]
for i in range(1, 16):
    fact.append(fact[-1] * i)
emit("INV_FACT", [Decimal(1) / f for f in fact])
