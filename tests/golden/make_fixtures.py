"""Writes tests/golden/reference_fixtures.json: the analytic operator matrices, stencil
weights and boundary-condition values that DiffEqOperators.jl's own test suite holds,
transcribed as numbers with their provenance (paths relative to the reference tree).

The reference is Julia and no Julia toolchain exists in this image, so nothing of the
reference is executed: every value below is read off the cited test file.  Each matrix
is the len x (len+2) matrix of the operator acting on a padded vector.

Run:  python tests/golden/make_fixtures.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def zeros(n, m):
    return [[0.0] * m for _ in range(n)]


def put(A, row, col, vals):
    """A[row, col:col+len(vals)] = vals, 1-based like the reference tests."""
    for k, v in enumerate(vals):
        A[row - 1][col - 1 + k] = float(v)


def fourth_deriv_fourth_approx(N):
    # test/DerivativeOperators/derivative_operators_interface.jl:8-21
    r1 = [3.5, -56 / 3, 42.5, -54.0, 251 / 6, -20.0, 5.5, -2 / 3]
    r2 = [2 / 3, -11 / 6, 0.0, 31 / 6, -22 / 3, 4.5, -4 / 3, 1 / 6]
    A = zeros(N, N + 2)
    put(A, 1, 1, r1)
    put(A, 2, 1, r2)
    put(A, N - 1, N - 5, r2[::-1])
    put(A, N, N - 5, r1[::-1])
    for i in range(3, N - 1):
        put(A, i, i - 2, [-1 / 6, 2.0, -13 / 2, 28 / 3, -13 / 2, 2.0, -1 / 6])
    return A


def second_deriv_fourth_approx(N):
    # test/DerivativeOperators/derivative_operators_interface.jl:23-31
    A = zeros(N, N + 2)
    put(A, 1, 1, [5 / 6, -15 / 12, -1 / 3, 7 / 6, -6 / 12, 5 / 60])
    put(A, N, N - 3, [1 / 12, -6 / 12, 14 / 12, -4 / 12, -15 / 12, 10 / 12])
    for i in range(2, N):
        put(A, i, i - 1, [-1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12])
    return A


def second_deriv_second_approx(N):
    # test/DerivativeOperators/derivative_operators_interface.jl:33-40
    A = zeros(N, N + 2)
    for i in range(1, N + 1):
        put(A, i, i, [1.0, -2.0, 1.0])
    return A


IRR_X = [0.0, 0.08, 0.1, 0.15, 0.19, 0.26, 0.29]
IRR_DX = [0.08, 0.02, 0.05, 0.04, 0.07, 0.03]


def ctr_irr():
    # test/DerivativeOperators/derivative_operators_interface.jl:48-100, used at :229-281
    one_two = zeros(5, 7)
    put(one_two, 1, 1, [-5.0 / 2.0, -75.0 / 2.0, 40.0])
    put(one_two, 2, 2, [-250.0 / 7.0, 30.0, 40.0 / 7.0])
    put(one_two, 3, 3, [-80.0 / 9.0, -5.0, 125.0 / 9.0])
    put(one_two, 4, 4, [-1225.0 / 77.0, 825.0 / 77.0, 400.0 / 77.0])
    put(one_two, 5, 5, [-30.0 / 7.0, -400.0 / 21.0, 70.0 / 3.0])
    two_two = zeros(5, 7)
    put(two_two, 1, 1, [250.0, -1250.0, 1000.0])
    put(two_two, 2, 2, [10000.0 / 7.0, -2000.0, 4000.0 / 7.0])
    put(two_two, 3, 3, [4000.0 / 9.0, -1000.0, 5000.0 / 9.0])
    put(two_two, 4, 4, [454.0 + 42.0 / 77.0, -5000.0 / 7.0, 20000.0 / 77.0])
    put(two_two, 5, 5, [6000.0 / 21.0, -20000.0 / 21.0, 2000.0 / 3.0])
    two_four = zeros(3, 7)
    put(two_four, 1, 1, [14.0 + 12012.0 / 13167.0, 1542.0 + 2736.0 / 13167.0,
                         -2288.0 - 11704.0 / 13167.0, 838.0 + 1254.0 / 13167.0,
                         -106.0 - 4298.0 / 13167.0])
    put(two_four, 2, 2, [-223.0 - 461.0 / 693.0, 847.0 + 154.0 / 693.0,
                         -1311.0 - 477.0 / 693.0, 699.0 + 593.0 / 693.0,
                         -11.0 - 502.0 / 693.0])
    put(two_four, 3, 3, [2.0 + 12166.0 / 13167.0, 538.0 + 12654.0 / 13167.0,
                         -912.0 - 9196.0 / 13167.0, 508.0 + 8664.0 / 13167.0,
                         -137.0 - 11121.0 / 13167.0])
    four_two = zeros(3, 7)
    put(four_two, 1, 1, [462000000.0 / 4389.0, -8550000000.0 / 4389.0, 11704000000.0 / 4389.0,
                         -5016000000.0 / 4389.0, 1400000000.0 / 4389.0])
    put(four_two, 2, 2, [200000000.0 / 231.0, -385000000.0 / 231.0, 360000000.0 / 231.0,
                         -200000000.0 / 231.0, 25000000.0 / 231.0])
    put(four_two, 3, 3, [770000000.0 / 4389.0, -3420000000.0 / 4389.0, 4180000000.0 / 4389.0,
                         -2850000000.0 / 4389.0, 1320000000 / 4389.0])
    return [
        dict(name="ctr_1_2_irr", d=1, a=2, rows="all", A=one_two,
             src="test/DerivativeOperators/derivative_operators_interface.jl:48-56,234-244"),
        dict(name="ctr_2_2_irr", d=2, a=2, rows="all", A=two_two,
             src="test/DerivativeOperators/derivative_operators_interface.jl:58-66,246-256"),
        dict(name="ctr_2_4_irr", d=2, a=4, rows="interior", A=two_four,
             src="test/DerivativeOperators/derivative_operators_interface.jl:68-80,258-268"),
        dict(name="ctr_4_2_irr", d=4, a=2, rows="interior", A=four_two,
             src="test/DerivativeOperators/derivative_operators_interface.jl:82-100,270-280"),
    ]


def upwind_uniform():
    # test/DerivativeOperators/upwind_operators_interface.jl:12-86 (uniform grid, dx = 1)
    one_one_pos = zeros(5, 7)
    for i in range(1, 6):
        put(one_one_pos, i, i + 1, [-1, 1])
    one_one_neg = zeros(5, 7)
    for i in range(1, 6):
        put(one_one_neg, i, i, [-1, 1])
    one_two_pos = zeros(5, 7)
    for i in range(1, 5):
        put(one_two_pos, i, i + 1, [-3 / 2, 2, -1 / 2])
    put(one_two_pos, 5, 5, [-1 / 2, 0, 1 / 2])
    one_two_neg = zeros(5, 7)
    put(one_two_neg, 1, 1, [-1 / 2, 0, 1 / 2])
    for i in range(2, 6):
        put(one_two_neg, i, i - 1, [1 / 2, -2, 3 / 2])
    two_two_pos = zeros(5, 7)
    for i in range(1, 4):
        put(two_two_pos, i, i + 1, [2, -5, 4, -1])
    put(two_two_pos, 4, 4, [1, -2, 1, 0])
    put(two_two_pos, 5, 4, [0, 1, -2, 1])
    two_two_neg = zeros(5, 7)
    put(two_two_neg, 1, 1, [1, -2, 1, 0])
    put(two_two_neg, 2, 1, [0, 1, -2, 1])
    for i in range(3, 6):
        put(two_two_neg, i, i - 2, [-1, 4, -5, 2])
    two_three_pos = zeros(7, 9)
    for i in range(1, 5):
        put(two_three_pos, i, i + 1, [35 / 12, -104 / 12, 114 / 12, -56 / 12, 11 / 12])
    put(two_three_pos, 5, 5, [11 / 12, -20 / 12, 6 / 12, 4 / 12, -1 / 12])
    put(two_three_pos, 6, 5, [-1 / 12, 16 / 12, -30 / 12, 16 / 12, -1 / 12])
    put(two_three_pos, 7, 5, [-1 / 12, 4 / 12, 6 / 12, -20 / 12, 11 / 12])
    two_three_neg = zeros(7, 9)
    put(two_three_neg, 1, 1, [11 / 12, -20 / 12, 6 / 12, 4 / 12, -1 / 12])
    put(two_three_neg, 2, 1, [-1 / 12, 16 / 12, -30 / 12, 16 / 12, -1 / 12])
    put(two_three_neg, 3, 1, [-1 / 12, 4 / 12, 6 / 12, -20 / 12, 11 / 12])
    for i in range(4, 8):
        put(two_three_neg, i, i - 3, [11 / 12, -56 / 12, 114 / 12, -104 / 12, 35 / 12])
    src = "test/DerivativeOperators/upwind_operators_interface.jl"
    out = []
    # (d, a, matrix, coeff, testset lines): the operator is UpwindDifference(d, a, 1.0, N,
    # coeff) and the expected matrix is coeff * analytic (the tests' `-1 * analytic...`).
    for name, d, a, A, c, lines in [
        ("up_1_1_pos", 1, 1, one_one_pos, 1.0, "12-18,152-179"),
        ("up_1_1_neg", 1, 1, one_one_neg, -1.0, "20-26,181-208"),
        ("up_1_2_pos", 1, 2, one_two_pos, 1.0, "28-35,210-241"),
        ("up_1_2_neg", 1, 2, one_two_neg, -1.0, "37-44,243-270"),
        ("up_2_2_pos", 2, 2, two_two_pos, 1.0, "46-54,272-299"),
        ("up_2_2_neg", 2, 2, two_two_neg, -1.0, "56-64,301-330"),
        ("up_2_3_pos", 2, 3, two_three_pos, 1.0, "66-75,332-359"),
        ("up_2_3_neg", 2, 3, two_three_neg, -1.0, "77-86,361-388"),
    ]:
        out.append(dict(name=name, d=d, a=a, dx=1.0, coeff=c,
                        A=[[c * v for v in row] for row in A], src=f"{src}:{lines}"))
    # scaled variants :570-784 -- expected = scale * analytic
    for name, d, a, base, dx, c, scale, lines in [
        ("up_1_2_pos_dx0.1", 1, 2, one_two_pos, 0.1, 1.0, 10.0, "570-598"),
        ("up_1_2_neg_dx0.1", 1, 2, one_two_neg, 0.1, -1.0, -10.0, "600-623"),
        ("up_2_2_pos_dx0.1", 2, 2, two_two_pos, 0.1, 1.0, 100.0, "625-648"),
        ("up_2_2_neg_dx0.1", 2, 2, two_two_neg, 0.1, -1.0, -100.0, "650-674"),
        ("up_2_2_pos_c4.56", 2, 2, two_two_pos, 1.0, 4.56, 4.56, "676-703"),
        ("up_2_2_neg_c4.56", 2, 2, two_two_neg, 1.0, -4.56, -4.56, "705-729"),
        ("up_2_2_pos_dx0.1_c4.56", 2, 2, two_two_pos, 0.1, 4.56, 4.56 * 100.0, "731-758"),
        ("up_2_2_neg_dx0.1_c4.56", 2, 2, two_two_neg, 0.1, -4.56, -4.56 * 100.0, "760-784"),
    ]:
        A = [[scale * v for v in row] for row in base]
        out.append(dict(name=name, d=d, a=a, dx=dx, coeff=c, A=A, src=f"{src}:{lines}"))
    return out


def upwind_irr():
    # test/DerivativeOperators/upwind_operators_interface.jl:92-150 (irregular grid)
    one_one_pos = zeros(5, 7)
    put(one_one_pos, 1, 2, [-50, 50])
    put(one_one_pos, 2, 3, [-20, 20])
    put(one_one_pos, 3, 4, [-25, 25])
    put(one_one_pos, 4, 5, [-100 / 7, 100 / 7])
    put(one_one_pos, 5, 6, [-100 / 3, 100 / 3])
    one_one_neg = zeros(5, 7)
    put(one_one_neg, 1, 1, [-25 / 2, 25 / 2])
    put(one_one_neg, 2, 2, [-50, 50])
    put(one_one_neg, 3, 3, [-20, 20])
    put(one_one_neg, 4, 4, [-25, 25])
    put(one_one_neg, 5, 5, [-100 / 7, 100 / 7])
    one_two_pos = zeros(5, 7)
    put(one_two_pos, 1, 2, [-450 / 7, 490 / 7, -40 / 7])
    put(one_two_pos, 2, 3, [-280 / 9, 405 / 9, -125 / 9])
    put(one_two_pos, 3, 4, [-2625 / 77, 3025 / 77, -400 / 77])
    put(one_two_pos, 4, 5, [-510 / 21, 1000 / 21, -490 / 21])
    put(one_two_pos, 5, 5, [-90 / 21, -400 / 21, 490 / 21])
    one_two_neg = zeros(5, 7)
    put(one_two_neg, 1, 1, [-5 / 2, -75 / 2, 80 / 2])
    put(one_two_neg, 2, 1, [5 / 2, -125 / 2, 120 / 2])
    put(one_two_neg, 3, 2, [250 / 7, -490 / 7, 240 / 7])
    put(one_two_neg, 4, 3, [80 / 9, -405 / 9, 325 / 9])
    put(one_two_neg, 5, 4, [1225 / 77, -3025 / 77, 1800 / 77])
    two_two_pos = zeros(5, 7)
    put(two_two_pos, 1, 2, [200000 / 77, -308000 / 77, 143000 / 77, -35000 / 77])
    put(two_two_pos, 2, 3, [27500 / 33, -75000 / 33, 55000 / 33, -7500 / 33])
    put(two_two_pos, 3, 4, [72500 / 77, -137500 / 77, 120000 / 77, -55000 / 77])
    put(two_two_pos, 4, 4, [42500 / 77, -71500 / 77, 40000 / 77, -11000 / 77])
    put(two_two_pos, 5, 4, [-10000 / 77, 44000 / 77, -100000 / 77, 66000 / 77])
    two_two_neg = zeros(5, 7)
    put(two_two_neg, 1, 1, [1050 / 7, -1250 / 7, -1400 / 7, 1600 / 7])
    put(two_two_neg, 2, 1, [350 / 7, 6250 / 7, -9800 / 7, 3200 / 7])
    put(two_two_neg, 3, 1, [-1400 / 7, 25000 / 7, -30800 / 7, 7200 / 7])
    put(two_two_neg, 4, 2, [-390000 / 231, 770000 / 231, -660000 / 231, 280000 / 231])
    put(two_two_neg, 5, 3, [-38500 / 77, 161000 / 77, -165000 / 77, 42500 / 77])
    src = "test/DerivativeOperators/upwind_operators_interface.jl"
    out = []
    for name, d, a, A, c, scale, lines in [
        ("up_1_1_pos_irr", 1, 1, one_one_pos, 1.0, 1.0, "92-100,390-418"),
        ("up_1_1_neg_irr", 1, 1, one_one_neg, -1.0, -1.0, "102-110,420-448"),
        ("up_1_2_pos_irr", 1, 2, one_two_pos, 1.0, 1.0, "112-120,450-478"),
        ("up_1_2_neg_irr", 1, 2, one_two_neg, -1.0, -1.0, "122-130,480-508"),
        ("up_2_2_pos_irr", 2, 2, two_two_pos, 1.0, 1.0, "132-140,510-538"),
        ("up_2_2_neg_irr", 2, 2, two_two_neg, -1.0, -1.0, "142-150,540-568"),
        ("up_2_2_pos_irr_c4.56", 2, 2, two_two_pos, 4.56, 4.56, "786-813"),
        ("up_2_2_neg_irr_c4.56", 2, 2, two_two_neg, -4.56, -4.56, "815-839"),
    ]:
        out.append(dict(name=name, d=d, a=a, dx=IRR_DX, coeff=c,
                        A=[[scale * v for v in row] for row in A], src=f"{src}:{lines}"))
    return out


def main():
    fx = {
        "about": "Values transcribed from DiffEqOperators.jl v4.45.0's test suite; see src of each entry.",
        "centered_uniform": [
            dict(name="ctr_4_4_N20", d=4, a=4, dx=1.0, N=20, A=fourth_deriv_fourth_approx(20),
                 src="test/DerivativeOperators/derivative_operators_interface.jl:8-21,121-133; "
                     "test/DerivativeOperators/convolutions.jl:17-30,64-81"),
            dict(name="ctr_2_4_N20", d=2, a=4, dx=1.0, N=20, A=second_deriv_fourth_approx(20),
                 src="test/DerivativeOperators/derivative_operators_interface.jl:23-31,134-143; "
                     "test/DerivativeOperators/convolutions.jl:32-40"),
            dict(name="ctr_2_2_N10", d=2, a=2, dx=1.0, N=10, A=second_deriv_second_approx(10),
                 src="test/DerivativeOperators/derivative_operators_interface.jl:33-40,284-295"),
            dict(name="ctr_2_2_N20", d=2, a=2, dx=1.0, N=20, A=second_deriv_second_approx(20),
                 src="test/DerivativeOperators/convolutions.jl:6-13,45-62"),
        ],
        "centered_irregular": dict(x=IRR_X, dx=IRR_DX, entries=ctr_irr()),
        # interior weights of centered stencils on a unit grid, d = 1..3, a in (2, 4)
        "centered_weights": dict(
            src="test/DerivativeOperators/derivative_operators_interface.jl:146-161",
            a2=[[-0.5, 0.0, 0.5], [1.0, -2.0, 1.0], [-1 / 2, 1.0, 0.0, -1.0, 1 / 2]],
            a4=[[1 / 12, -2 / 3, 0.0, 2 / 3, -1 / 12], [-1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12],
                [1 / 8, -1.0, 13 / 8, 0.0, -13 / 8, 1.0, -1 / 8]],
        ),
        # forward (offside 0) upwind weights on a unit grid: [a-1][d in (1, 3)]
        "upwind_weights": dict(
            src="test/DerivativeOperators/derivative_operators_interface.jl:203-214",
            a1=[[-1.0, 1.0], [-1.0, 3.0, -3.0, 1.0]],
            a2=[[-3 / 2, 2.0, -1 / 2], [-5 / 2, 9.0, -12.0, 7.0, -3 / 2]],
        ),
        "upwind_uniform": upwind_uniform(),
        "upwind_irregular": upwind_irr(),
        # 3rd-order Robin ghosts, u = 1:10, (alpha, beta, gamma) = (1, 6, 10), dx = 1
        "robin3": dict(src="test/DerivativeOperators/robin.jl:92-100", u=list(range(1, 11)),
                       l=(1.0, 6.0, 10.0), r=(1.0, 6.0, 10.0), dx=1.0, order=3,
                       ghost_l=-4 / 10, ghost_r=125 / 12),
    }
    path = os.path.join(HERE, "reference_fixtures.json")
    with open(path, "w") as f:
        json.dump(fx, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
