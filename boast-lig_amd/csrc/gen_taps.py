#!/usr/bin/env python3
"""Write lc_taps.h: the sym8 filter taps of the libconv operators.

The decimal values below are the published Daubechies symlet-8 tables used
by BigDFT (scaling filter, magic filter, first/second-derivative filters).
They are mathematical constants, listed with the same digits as the
reference tables (BigDFT/WaveletFilters.rb:25-41 SYM8_LP, :43-72 SYM8_D1,
:74-103 SYM8_D2, :105-121 SYM8_MF).  Each filter is emitted as an X-macro so
that C (the oracle) and HIP (the kernels) instantiate it in double and float
with the compiler's correctly rounded decimal->binary conversion (a float
tap is rounded once from the decimal string, never through double).

Filters and their index ranges (tap l of filter F sits at F[l - lo]):
  MD  magic filter          l in [-8, 7]   = SYM8_MF reversed, centre 8
                                            (BigDFT/LibconvBase.rb:239,
                                             MagicFilter1d.rb:30)
  IMD inverse magic filter  l in [-7, 8]   = SYM8_MF, centre 7
                                            (BigDFT/LibconvBase.rb:245)
  D2  kinetic (2nd deriv.)  l in [-14, 14] = SYM8_D2, centre 14
                                            (BigDFT/Kinetic1d.rb:33)
  D1  gradient (1st deriv.) l in [-14, 14] = SYM8_D1, centre 14
  LP  sym8 low-pass k in [0, 15]; HP[k] = (-1)^(k+1) LP[15-k]
      (BigDFT/GenericConvolution2.rb:306-325)
  CLP/CHP  combined LP (x) MF filters of the s1tor/rtos1 operators,
      computed exactly here in decimal and truncated to 30 decimal places
      (BigDFT/GenericConvolution2.rb:336-384); one pair for the periodic
      centre (8) and one for the grow/shrink centre (7).

Run:  python3 gen_taps.py  (rewrites lc_taps.h next to this file).
"""
import decimal
import os

decimal.getcontext().prec = 200

SYM8_LP = """
-0.003382415951005002595457699235919135128243170037709254473102252636e0
-0.000542132331800010689347836944017347858824603460398208422132403418e0
0.03169508781152599143142570752391597485717789804058994936458784274e0
0.00760748732497660819192100801219238392349477628276637389848185158e0
-0.1432942383512726628440955443992739699342280444601654901737376736e0
-0.0612733590678110778430467690736584126556319554926068967930152121e0
0.481359651259053391589568580264984525838564867578278365157821169e0
0.777185751699628028624333566639210809121328610967576793053795894e0
0.3644418948361789367595594323842960136510635625353935885398873954e0
-0.0519458381078818007357107290625634285296352747007956649813654351e0
-0.02721902991710348632196412421624687935352227243191326268168750639e0
0.0491371796737302867869109881551654429015527772546123162293717851e0
0.003808752013894489463071922223049694342708315178928653470691661633e0
-0.01495225833706219911849029822472598879079395753054027354116270039e0
-0.0003029205147241330812639124399571849886852187149285126161207667492e0
0.001889950332767689184274432603245581173345564367859597144366089742e0
""".split()

SYM8_MF = """
2.727344929119796596577153130172279928753156083478007592916189e-6
-0.00005185986881173432922848639136911486874128383207856851537290756e0
0.0004944322768868991919228225947675097243713535139015057534937039e0
-0.003441281444934938572808815096868218611693146749743898558729002e0
0.013372634148547947527334234670132209969450042970835150850985441e0
-0.02103025160930381434955489412839065067238766803542905035107527e0
-0.06048952891969835160028346363993374032323310579329060307398682e0
0.99404156978340039931786167125707312696826875197808439082533262e0
0.06126258958312079821953805969111734845421975622016765542503018e0
0.02373821463724942397566389712597274534984899296281082155336152e0
-0.00942047030201080385922711540948195074607340327611386232053071e0
0.001747237136729939034494478127498529416920979083511241686685502e0
-0.000301580381326904631671637038261698791032841826554967601640246e0
0.000087629844762105595646891618941163970464425424613649618138794e0
-0.00001290557201342060969516786758559028264129757158423127355882943e0
8.4334247333529341094733325815816232969177478728797427310317e-7
""".split()

# half of the symmetric second-derivative filter: taps 0..14 (tap -l = tap l)
SYM8_D2_HALF = """
-3.553692289913190194129680937424076424953761573153458649028535991e0
2.219146593891116389879454640457705235616912723375223389717990369e0
-0.6156141465570069496314853948700762181701834843585057391570819072e0
0.2371780582153805636239247475762231531845465033753480080361580487e0
-0.0822663999742123340987663520631668159519522793487327848931628121e0
0.02207029188482255523789911295638968408882184950053649882554342682e0
-0.004097656893426338238993270511883154847227992660459896521032712738e0
0.000451679202875022353494800376397584960775566033384647821996975667e0
-0.00002398228524507599670405555359023134858523138758236007644059252025e0
2.09042349529203659579228894473609842180443353287738050051136428e-6
-3.723076304736927584879149697304408803274215476409064356088861068e-7
-1.058570554967414703734941322870242653300359680996853899470022889e-8
-5.813879830282540547959250667464887365484326600877085719823602219e-11
2.708004936263194382698566890376475760883688432060067551662628469e-13
-6.924474940639200152025730585882115968579663856946045039077208122e-18
""".split()

# half of the antisymmetric first-derivative filter: taps 1..14 (tap 0 = 0,
# tap -l = -tap l)
SYM8_D1_HALF = """
0.8834460460908270942785855643941098713157598590915435134511547387e0
-0.3032593514765938346887962440931660566677778365072429185627922116e0
0.1063640682894442760934532284702982844094052663346279226663978027e0
-0.03129014783948023634381563995259325360348542944538434049761056972e0
0.006958379116450707495020407536758241511060163284033091893516844663e0
-0.001031530213375445369097964780217146384828800932615832761481602166e0
0.00007667706908380351933901775234801819926013328778633133146782374641e0
2.451992111053665419191564123434446284949933594307826864375530275e-7
3.993810456408053712133667496514167576517175934074503260991290434e-8
-7.207948238588481597101904067961163265788433149563519306241683057e-8
-9.697184925637300947553068758362746388131353737840306778881714191e-10
-7.252206916665149851135591529994603533072998149301256702737646004e-13
1.240078536096648534547438612731567439237828505396469976633162949e-14
-1.585464751677102510097179308121926878193105545088903963493188933e-19
""".split()


def neg(s):
    """Negate a decimal string exactly (textual sign flip)."""
    return s[1:] if s.startswith('-') else '-' + s


def d2_full():
    half = SYM8_D2_HALF
    return [half[abs(l)] for l in range(-14, 15)]


def d1_full():
    out = []
    for l in range(-14, 15):
        if l == 0:
            out.append('0e0')
        elif l > 0:
            out.append(SYM8_D1_HALF[l - 1])
        else:
            out.append(neg(SYM8_D1_HALF[-l - 1]))
    return out


def hp_from_lp(lp):
    n = len(lp)
    return [lp[n - 1 - k] if k % 2 == 1 else neg(lp[n - 1 - k]) for k in range(n)]


def combined(lp_str, cv_str, cv_center):
    """Exact decimal convolution of a wavelet filter with the magic filter.

    Restates BigDFT/GenericConvolution2.rb:336-384: the bounds of the
    product filter are widened so that its lower bound is even and its upper
    bound odd, and every tap is truncated to 30 decimal places.
    Returns (taps as decimal strings, centre).
    """
    D = decimal.Decimal
    lp = [D(s) for s in lp_str]
    cv = [D(s) for s in cv_str]
    n_lp, n_cv = len(lp), len(cv)
    lp_c = n_lp // 2
    cv_lo, cv_up = -cv_center, n_cv - cv_center - 1
    lp_lo, lp_up = -lp_c, n_lp - lp_c - 1
    lo, up = cv_lo - lp_up, cv_up - lp_lo
    if lo % 2 != 0:
        lo -= 1
    if up % 2 == 0:
        up += 1
    taps = []
    q = D(1).scaleb(-30)
    for i in range(lo, up + 1):
        s = D(0)
        for j in range(lp_lo, lp_up + 1):
            k = i - j - 1 + cv_center
            if 0 <= k < n_cv:
                s += lp[j + lp_c] * cv[k]
        taps.append(s.quantize(q, rounding=decimal.ROUND_DOWN))
    return ['%s' % t for t in taps], -lo


def poisson_table(nord):
    """Finite-difference first-derivative weights of order nord (the Poisson
    nabla filters, BigDFT/Poisson2.rb generate_filter): row k in [-m, m]
    (m = nord / 2) holds the weights c(j, k), j in [-m, m], of the
    (nord + 1)-point stencil on the nodes j - k around the output point, so
    that sum_j c(j, k) f(j - k) = f'(0) for every polynomial f of degree
    <= nord.  Row 0 is the centred stencil; rows k < 0 serve the first m
    outputs of a non-periodic line (output i uses row i - m and inputs
    0 .. 2m), rows k > 0 the last m (GenericConvolution2.rb:237-294).  The
    weights are computed exactly (Lagrange) and emitted as quotients of
    integers, so each rounds once to the storage precision."""
    from fractions import Fraction
    m = nord // 2
    rows = []
    for k in range(-m, m + 1):
        nodes = [j - k for j in range(-m, m + 1)]
        w = []
        for a in nodes:
            others = [b for b in nodes if b != a]
            den = Fraction(1)
            for b in others:
                den *= a - b
            num = Fraction(0)  # derivative at 0 of prod_{b != a} (x - b)
            for c in others:
                p = Fraction(1)
                for b in others:
                    if b != c:
                        p *= -b
                num += p
            w.append(num / den)
        rows.append(w)
    return rows


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    lp = SYM8_LP
    hp = hp_from_lp(lp)
    md = list(reversed(SYM8_MF))
    imd = list(SYM8_MF)
    filters = [
        ('MD', md, -8, 'magic filter, l in [-8, 7]'),
        ('IMD', imd, -7, 'inverse (transposed) magic filter, l in [-7, 8]'),
        ('D2', d2_full(), -14, 'second-derivative (kinetic) filter, l in [-14, 14]'),
        ('D1', d1_full(), -14, 'first-derivative filter, l in [-14, 14]'),
        ('LP', lp, 0, 'sym8 low-pass scaling filter, k in [0, 15]'),
        ('HP', hp, 0, 'sym8 high-pass filter HP[k] = (-1)^(k+1) LP[15-k]'),
    ]
    for cen, tag in ((8, 'P'), (7, 'G')):
        clp, clc = combined(lp, SYM8_MF, cen)
        chp, chc = combined(hp, SYM8_MF, cen)
        assert clc == chc and len(clp) == len(chp) == 32
        what = 'periodic' if tag == 'P' else 'grow/shrink'
        filters.append(('CLP' + tag, clp, -clc,
                        'LP (x) MF combined filter, %s centre %d' % (what, clc)))
        filters.append(('CHP' + tag, chp, -chc,
                        'HP (x) MF combined filter, %s centre %d' % (what, chc)))

    lines = [
        '/* lc_taps.h -- generated by gen_taps.py from the decimal sym8 tables;',
        ' * do not edit.  LC_TAPS_<F>(X) expands X(value) for every tap of filter',
        ' * F in increasing tap index; LC_<F>_LO is the index of the first tap and',
        ' * LC_<F>_N the number of taps. */',
        '#ifndef LC_TAPS_H',
        '#define LC_TAPS_H',
        '',
    ]
    for name, taps, lo, doc in filters:
        lines.append('/* %s */' % doc)
        lines.append('#define LC_%s_LO (%d)' % (name, lo))
        lines.append('#define LC_%s_N %d' % (name, len(taps)))
        lines.append('#define LC_TAPS_%s(X) \\' % name)
        for i, t in enumerate(taps):
            end = ' \\' if i + 1 < len(taps) else ''
            lines.append('  X(%s)%s' % (t, end))
        lines.append('')
    lines.append('/* Poisson nabla filters: LC_POISSON_<nord>(X) expands X(p, q) for the')
    lines.append(' * weight p / q of row k, tap j, rows k = -m .. m outer, taps j = -m .. m')
    lines.append(' * inner (m = nord / 2); see poisson_table() in gen_taps.py. */')
    for nord in (2, 4, 6, 8, 16):
        rows = poisson_table(nord)
        lines.append('#define LC_POISSON_%d(X) \\' % nord)
        flat = [w for r in rows for w in r]
        for i, w in enumerate(flat):
            end = ' \\' if i + 1 < len(flat) else ''
            lines.append('  X(%d, %d)%s' % (w.numerator, w.denominator, end))
        lines.append('')
    lines.append('#endif /* LC_TAPS_H */')
    with open(os.path.join(here, 'lc_taps.h'), 'w') as f:
        f.write('\n'.join(lines) + '\n')


if __name__ == '__main__':
    main()
