"""Writes bv_rgamma_coef.h: Taylor coefficients a_k of 1/Gamma(1+z) at z=0,
computed with mpmath at 50 digits.  Used by the Temme series for K_nu
(|mu| <= 1/2), where 1/Gamma(1+mu), 1/Gamma(1-mu) and the two Temme
combinations gamma1, gamma2 are evaluated from this series without
cancellation.  Run: python gen_rgamma.py > bv_rgamma_coef.h"""
import mpmath as mp

mp.mp.dps = 50
N = 26
coef = mp.taylor(lambda z: 1 / mp.gamma(1 + z), 0, N - 1)
print("/* Generated by gen_rgamma.py: Taylor coefficients of 1/Gamma(1+z) at 0. */")
print("#ifndef BV_RGAMMA_COEF_H")
print("#define BV_RGAMMA_COEF_H")
print("#define BV_RGAMMA_NCOEF %d" % N)
print("#define BV_RGAMMA_COEFS { \\")
for k, c in enumerate(coef):
    print("    %s%s \\" % (mp.nstr(c, 20, min_fixed=0, max_fixed=0), "," if k < N - 1 else ""))
print("}")
print("#endif")
