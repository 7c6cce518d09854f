/*
 * imfcs_oracle.c -- CPU restatement of the Imaging_FCS 1.52 Java CPU path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  See imfcs_oracle.h.
 *
 * Sources followed (IFCS = Imaging_FCS_1_52.java):
 *   tables            IFCS:7343, 7391-7424
 *   intensity         IFCS:9651-9678 (getIntensity), 9822-9855 (trace)
 *   bleach            IFCS:9734-9769, 10953-11004, 13088-13117
 *   blocking          IFCS:9217-9422 (blockTransform)
 *   correlation       IFCS:9425-9648 (calculateCF, incl. GLS covariance)
 *   model             IFCS:13131-13408 (FCS_3p), 13670-13702 (obsvol)
 *   fits              IFCS:10288-10471 (standardFit), 10622-10789 (GLSFit),
 *                     13569-13729 (GLS_fitFunction, dataTransform)
 *   ROI loop          IFCS:8437-8544 (correlateROI, CPU branch)
 *
 * The least-squares optimizer is the Levenberg-Marquardt method of
 * Apache commons-math3 3.6.1 (LevenbergMarquardtOptimizer), which is a
 * translation of MINPACK lmder (More, Garbow, Hillstrom 1980).  That
 * library is not part of the reference tree; its published algorithm is
 * restated here with its default settings (initialStepBoundFactor 100,
 * cost/parameter/orthogonality relative tolerances 1e-10, QR ranking
 * threshold = Double.MIN_NORMAL).  Its call sites are IFCS:10329-10340
 * (standardFit), 10660-10666 (GLSFit) and 10977-10984 (PolynomFit).
 */
#include "imfcs_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* a1: correlator tables (IFCS:7343, 7391-7424)                        */
/* ------------------------------------------------------------------ */
int orc_tables(int p, int q, int num, double frametime,
               int *lag, double *samp, double *mtab, double *lagtime)
{
    int base = p, hbase = p / 2, lagnum = q;
    int chanum = (int)(p + (q - 1) * p / 2 + 1);
    if (chanum > ORC_MAXCHAN) return -1;
    for (int x = 0; x <= hbase; x++) {
        lag[x] = x;
        lagtime[x] = lag[x] * frametime;
    }
    for (int x = 1; x <= lagnum; x++) {
        for (int y = 1; y <= hbase; y++) {
            lag[x * hbase + y] = (int)(pow(2, x - 1) * y + (base / 4) * pow(2, x));
            lagtime[x * hbase + y] = lag[x * hbase + y] * frametime;
        }
    }
    for (int x = 0; x <= base; x++) samp[x] = 1;
    for (int x = 2; x <= lagnum; x++)
        for (int y = 1; y <= hbase; y++) samp[x * hbase + y] = pow(2, x - 1);
    for (int x = 0; x <= chanum - 1; x++)
        mtab[x] = (int)floor((num - lag[x]) / samp[x]);
    return chanum;
}

/* ------------------------------------------------------------------ */
/* a3: getIntensity binning + background (IFCS:9668-9678)             */
/* ------------------------------------------------------------------ */
void orc_get_intensity(const uint16_t *stack, int W, int H, int T,
                       int px, int py, int bx, int by, int bckg,
                       int initialframe, int finalframe, double *intdat)
{
    (void)T;
    int num = finalframe - initialframe + 1;
    const size_t plane = (size_t)W * H;
    intdat[0] = 0;
    for (int x = 1; x <= num; x++) {
        const uint16_t *fr = stack + (size_t)(initialframe + x - 2) * plane;
        double acc = 0;
        for (int i = 0; i < bx; i++)
            for (int k = 0; k < by; k++)
                acc += (int)fr[(size_t)(py + k) * W + (px + i)] - bckg;
        intdat[x] = acc;
    }
}

/* a4: calcIntensityTrace (IFCS:9822-9855), integer sums and integer
 * division by ave exactly as the Java int arithmetic does. */
void orc_intensity_trace(const uint16_t *stack, int W, int H, int T,
                         int px, int py, int bx, int by, int bckg,
                         int initialframe, int finalframe, int nopit,
                         double frametime, double *trace, double *intTime)
{
    (void)T;
    int ave = (finalframe - initialframe + 1) / nopit;
    const size_t plane = (size_t)W * H;
    for (int x = 0; x < nopit; x++) {
        int sum1 = 0;
        for (int i = 0; i < bx; i++)
            for (int k = 0; k < by; k++)
                for (int y = initialframe + x * ave; y <= initialframe + (x + 1) * ave - 1; y++)
                    sum1 += (int)stack[(size_t)(y - 1) * plane + (size_t)(py + k) * W + (px + i)] - bckg;
        intTime[x] = frametime * (x + 0.5) * ave;
        trace[x] = (double)(sum1 / ave);
    }
}

/* ------------------------------------------------------------------ */
/* Levenberg-Marquardt: commons-math3 3.6.1 LevenbergMarquardtOptimizer */
/* (a translation of MINPACK lmder/lmpar/qrsolv).                      */
/* Matrices: wj[i*n + j] is row i (observation), column j (parameter).  */
/* ------------------------------------------------------------------ */
#define LM_INIT_STEP 100.0
#define LM_COST_TOL 1e-10
#define LM_PAR_TOL 1e-10
#define LM_ORTHO_TOL 1e-10
#define LM_TWO_EPS (2 * 1.1102230246251565e-16)
#define LM_SAFE_MIN DBL_MIN

typedef struct {
    int m, n, rank;
    double *wj;      /* weighted jacobian, then R + Householder vectors */
    int *perm;
    double *diagR, *jacNorm, *beta;
} lm_qr;

/* -(W^1/2 J) QR with column pivoting (qrDecomposition). Returns 0 or -1. */
static int lm_qr_decompose(lm_qr *q, int solvedCols)
{
    int nR = q->m, nC = q->n;
    double *a = q->wj;
    for (int k = 0; k < nC; ++k) {
        q->perm[k] = k;
        double norm2 = 0;
        for (int i = 0; i < nR; ++i) {
            double akk = a[i * nC + k];
            norm2 += akk * akk;
        }
        q->jacNorm[k] = sqrt(norm2);
    }
    for (int k = 0; k < nC; ++k) {
        int nextColumn = -1;
        double ak2 = -INFINITY;
        for (int i = k; i < nC; ++i) {
            double norm2 = 0;
            for (int j = k; j < nR; ++j) {
                double aki = a[j * nC + q->perm[i]];
                norm2 += aki * aki;
            }
            if (isinf(norm2) || isnan(norm2)) return -1;
            if (norm2 > ak2) {
                nextColumn = i;
                ak2 = norm2;
            }
        }
        if (ak2 <= LM_SAFE_MIN) {
            q->rank = k;
            return 0;
        }
        int pk = q->perm[nextColumn];
        q->perm[nextColumn] = q->perm[k];
        q->perm[k] = pk;
        double akk = a[k * nC + pk];
        double alpha = (akk > 0) ? -sqrt(ak2) : sqrt(ak2);
        double betak = 1.0 / (ak2 - akk * alpha);
        q->beta[pk] = betak;
        q->diagR[pk] = alpha;
        a[k * nC + pk] -= alpha;
        for (int dk = nC - 1 - k; dk > 0; --dk) {
            double gamma = 0;
            int col = q->perm[k + dk];
            for (int j = k; j < nR; ++j) gamma += a[j * nC + pk] * a[j * nC + col];
            gamma *= betak;
            for (int j = k; j < nR; ++j) a[j * nC + col] -= gamma * a[j * nC + pk];
        }
    }
    q->rank = solvedCols;
    return 0;
}

static void lm_qty(double *y, const lm_qr *q)
{
    int nR = q->m, nC = q->n;
    for (int k = 0; k < nC; ++k) {
        int pk = q->perm[k];
        double gamma = 0;
        for (int i = k; i < nR; ++i) gamma += q->wj[i * nC + pk] * y[i];
        gamma *= q->beta[pk];
        for (int i = k; i < nR; ++i) y[i] -= gamma * q->wj[i * nC + pk];
    }
}

/* determineLMDirection (MINPACK qrsolv). */
static void lm_direction(const double *qy, const double *diag, double *lmDiag,
                         lm_qr *q, int solvedCols, double *work, double *lmDir)
{
    int nC = q->n;
    double *a = q->wj;
    const int *perm = q->perm;
    for (int j = 0; j < solvedCols; ++j) {
        int pj = perm[j];
        for (int i = j + 1; i < solvedCols; ++i) a[i * nC + pj] = a[j * nC + perm[i]];
        lmDir[j] = q->diagR[pj];
        work[j] = qy[j];
    }
    for (int j = 0; j < solvedCols; ++j) {
        int pj = perm[j];
        double dpj = diag[pj];
        if (dpj != 0) {
            for (int i = j + 1; i < nC; i++) lmDiag[i] = 0;
        }
        lmDiag[j] = dpj;
        double qtbpj = 0;
        for (int k = j; k < solvedCols; ++k) {
            int pk = perm[k];
            if (lmDiag[k] != 0) {
                double sn, cs;
                double rkk = a[k * nC + pk];
                if (fabs(rkk) < fabs(lmDiag[k])) {
                    double cotan = rkk / lmDiag[k];
                    sn = 1.0 / sqrt(1.0 + cotan * cotan);
                    cs = sn * cotan;
                } else {
                    double tn = lmDiag[k] / rkk;
                    cs = 1.0 / sqrt(1.0 + tn * tn);
                    sn = cs * tn;
                }
                a[k * nC + pk] = cs * rkk + sn * lmDiag[k];
                double temp = cs * work[k] + sn * qtbpj;
                qtbpj = -sn * work[k] + cs * qtbpj;
                work[k] = temp;
                for (int i = k + 1; i < solvedCols; ++i) {
                    double rik = a[i * nC + pk];
                    double temp2 = cs * rik + sn * lmDiag[i];
                    lmDiag[i] = -sn * rik + cs * lmDiag[i];
                    a[i * nC + pk] = temp2;
                }
            }
        }
        lmDiag[j] = a[j * nC + perm[j]];
        a[j * nC + perm[j]] = lmDir[j];
    }
    int nSing = solvedCols;
    for (int j = 0; j < solvedCols; ++j) {
        if ((lmDiag[j] == 0) && (nSing == solvedCols)) nSing = j;
        if (nSing < solvedCols) work[j] = 0;
    }
    if (nSing > 0) {
        for (int j = nSing - 1; j >= 0; --j) {
            int pj = perm[j];
            double sum = 0;
            for (int i = j + 1; i < nSing; ++i) sum += a[i * nC + pj] * work[i];
            work[j] = (work[j] - sum) / lmDiag[j];
        }
    }
    for (int j = 0; j < nC; ++j) lmDir[perm[j]] = work[j];
}

/* determineLMParameter (MINPACK lmpar). */
static double lm_parameter(const double *qy, double delta, const double *diag,
                           lm_qr *q, int solvedCols, double *work1,
                           double *work2, double *work3, double *lmDir,
                           double lmPar)
{
    int nC = q->n, rank = q->rank;
    const double *a = q->wj;
    const int *perm = q->perm;
    for (int j = 0; j < rank; ++j) lmDir[perm[j]] = qy[j];
    for (int j = rank; j < nC; ++j) lmDir[perm[j]] = 0;
    for (int k = rank - 1; k >= 0; --k) {
        int pk = perm[k];
        double ypk = lmDir[pk] / q->diagR[pk];
        for (int i = 0; i < k; ++i) lmDir[perm[i]] -= ypk * a[i * nC + pk];
        lmDir[pk] = ypk;
    }
    double dxNorm = 0;
    for (int j = 0; j < solvedCols; ++j) {
        int pj = perm[j];
        double s = diag[pj] * lmDir[pj];
        work1[pj] = s;
        dxNorm += s * s;
    }
    dxNorm = sqrt(dxNorm);
    double fp = dxNorm - delta;
    if (fp <= 0.1 * delta) return 0;

    double sum2, parl = 0;
    if (rank == solvedCols) {
        for (int j = 0; j < solvedCols; ++j) {
            int pj = perm[j];
            work1[pj] *= diag[pj] / dxNorm;
        }
        sum2 = 0;
        for (int j = 0; j < solvedCols; ++j) {
            int pj = perm[j];
            double sum = 0;
            for (int i = 0; i < j; ++i) sum += a[i * nC + pj] * work1[perm[i]];
            double s = (work1[pj] - sum) / q->diagR[pj];
            work1[pj] = s;
            sum2 += s * s;
        }
        parl = fp / (delta * sum2);
    }
    sum2 = 0;
    for (int j = 0; j < solvedCols; ++j) {
        int pj = perm[j];
        double sum = 0;
        for (int i = 0; i <= j; ++i) sum += a[i * nC + pj] * qy[i];
        sum /= diag[pj];
        sum2 += sum * sum;
    }
    double gNorm = sqrt(sum2);
    double paru = gNorm / delta;
    if (paru == 0) paru = LM_SAFE_MIN / fmin(delta, 0.1);
    lmPar = fmin(paru, fmax(lmPar, parl));
    if (lmPar == 0) lmPar = gNorm / dxNorm;

    for (int countdown = 10; countdown >= 0; --countdown) {
        if (lmPar == 0) lmPar = fmax(LM_SAFE_MIN, 0.001 * paru);
        double sPar = sqrt(lmPar);
        for (int j = 0; j < solvedCols; ++j) {
            int pj = perm[j];
            work1[pj] = sPar * diag[pj];
        }
        lm_direction(qy, work1, work2, q, solvedCols, work3, lmDir);
        dxNorm = 0;
        for (int j = 0; j < solvedCols; ++j) {
            int pj = perm[j];
            double s = diag[pj] * lmDir[pj];
            work3[pj] = s;
            dxNorm += s * s;
        }
        dxNorm = sqrt(dxNorm);
        double previousFP = fp;
        fp = dxNorm - delta;
        if (fabs(fp) <= 0.1 * delta || (parl == 0 && fp <= previousFP && previousFP < 0))
            return lmPar;
        for (int j = 0; j < solvedCols; ++j) {
            int pj = perm[j];
            work1[pj] = work3[pj] * diag[pj] / dxNorm;
        }
        for (int j = 0; j < solvedCols; ++j) {
            int pj = perm[j];
            work1[pj] /= work2[j];
            double tmp = work1[pj];
            for (int i = j + 1; i < solvedCols; ++i) work1[perm[i]] -= a[i * nC + pj] * tmp;
        }
        sum2 = 0;
        for (int j = 0; j < solvedCols; ++j) {
            double s = work1[perm[j]];
            sum2 += s * s;
        }
        double correction = fp / (delta * sum2);
        if (fp > 0) parl = fmax(parl, lmPar);
        else if (fp < 0) paru = fmin(paru, lmPar);
        lmPar = fmax(parl, lmPar + correction);
    }
    return lmPar;
}

/* Evaluate weighted residuals r = sqrt(w) (target - f) and the weighted
 * jacobian sqrt(w) J (LeastSquaresFactory.weightMatrix semantics). */
static double lm_eval(int m, int n, const double *x, const double *target,
                      const double *sqw, orc_model_fn fn, void *ctx,
                      double *val, double *jac, double *res)
{
    fn(ctx, x, val, jac);
    double cost = 0;
    for (int i = 0; i < m; i++) {
        res[i] = sqw[i] * (target[i] - val[i]);
        cost += res[i] * res[i];
        for (int j = 0; j < n; j++) jac[i * n + j] *= sqw[i];
    }
    return sqrt(cost);
}

int orc_lm(int m, int n, double *x, const double *target,
           const double *weights, orc_model_fn fn, void *ctx,
           long maxEval, long maxIter, double *wresid_out,
           int *nevals, int *niters)
{
    int status = ORC_LM_OK;
    int nR = m, nC = n, solvedCols = (m < n) ? m : n;
    double *buf = (double *)calloc((size_t)(8 * nC + 6 * nR + 3 * nR * nC + 4), sizeof(double));
    int *perm = (int *)calloc((size_t)nC, sizeof(int));
    double *lmDir = buf, *diag = lmDir + nC, *oldX = diag + nC, *work1 = oldX + nC;
    double *work2 = work1 + nC, *work3 = work2 + nC, *diagR = work3 + nC;
    double *jacNorm = diagR + nC;
    double *betav = (double *)calloc((size_t)nC, sizeof(double));
    double *sqw = jacNorm + nC, *qtf = sqw + nR, *resA = qtf + nR, *resB = resA + nR;
    double *val = resB + nR, *resSave = val + nR;
    double *jac = resSave + nR, *jacPrev = jac + nR * nC, *wj = jacPrev + nR * nC;
    double *pt = (double *)malloc(sizeof(double) * (size_t)nC);
    long evalCount = 0, iterCount = 0;
    double lmPar = 0, delta = 0, xNorm = 0;
    for (int i = 0; i < nR; i++) sqw[i] = weights ? sqrt(weights[i]) : 1.0;
    for (int j = 0; j < nC; j++) pt[j] = x[j];

    lm_qr q = {nR, nC, 0, wj, perm, diagR, jacNorm, betav};

    /* "current" evaluation: residuals in curRes, jacobian in jac */
    evalCount++;
    double *curRes = resA, *oldRes = resB;
    double currentCost = lm_eval(nR, nC, pt, target, sqw, fn, ctx, val, jac, curRes);
    int firstIteration = 1;
    for (;;) {
        if (++iterCount > maxIter) { status = ORC_LM_TOO_MANY_ITER; goto done; }
        /* previous = current: keep its jacobian and residuals */
        memcpy(jacPrev, jac, sizeof(double) * nR * nC);
        memcpy(resSave, curRes, sizeof(double) * nR);
        for (int i = 0; i < nR * nC; i++) wj[i] = -jac[i];
        if (lm_qr_decompose(&q, solvedCols) != 0) { status = ORC_LM_QR_FAIL; goto done; }
        double *weightedResidual = curRes;
        for (int i = 0; i < nR; i++) qtf[i] = weightedResidual[i];
        lm_qty(qtf, &q);
        for (int k = 0; k < solvedCols; ++k) {
            int pk = perm[k];
            wj[k * nC + pk] = diagR[pk];
        }
        if (firstIteration) {
            xNorm = 0;
            for (int k = 0; k < nC; ++k) {
                double dk = jacNorm[k];
                if (dk == 0) dk = 1.0;
                double xk = dk * pt[k];
                xNorm += xk * xk;
                diag[k] = dk;
            }
            xNorm = sqrt(xNorm);
            delta = (xNorm == 0) ? LM_INIT_STEP : (LM_INIT_STEP * xNorm);
        }
        double maxCosine = 0;
        if (currentCost != 0) {
            for (int j = 0; j < solvedCols; ++j) {
                int pj = perm[j];
                double s = jacNorm[pj];
                if (s != 0) {
                    double sum = 0;
                    for (int i = 0; i <= j; ++i) sum += wj[i * nC + pj] * qtf[i];
                    maxCosine = fmax(maxCosine, fabs(sum) / (s * currentCost));
                }
            }
        }
        if (maxCosine <= LM_ORTHO_TOL) goto done;
        for (int j = 0; j < nC; ++j) diag[j] = fmax(diag[j], jacNorm[j]);

        for (double ratio = 0; ratio < 1.0e-4;) {
            for (int j = 0; j < solvedCols; ++j) {
                int pj = perm[j];
                oldX[pj] = pt[pj];
            }
            double previousCost = currentCost;
            double *tmpVec = weightedResidual;
            weightedResidual = oldRes;
            oldRes = tmpVec;

            lmPar = lm_parameter(qtf, delta, diag, &q, solvedCols, work1, work2, work3, lmDir, lmPar);
            double lmNorm = 0;
            for (int j = 0; j < solvedCols; ++j) {
                int pj = perm[j];
                lmDir[pj] = -lmDir[pj];
                pt[pj] = oldX[pj] + lmDir[pj];
                double s = diag[pj] * lmDir[pj];
                lmNorm += s * s;
            }
            lmNorm = sqrt(lmNorm);
            if (firstIteration) delta = fmin(delta, lmNorm);

            if (++evalCount > maxEval) { status = ORC_LM_TOO_MANY_EVAL; goto done; }
            currentCost = lm_eval(nR, nC, pt, target, sqw, fn, ctx, val, jac, weightedResidual);
            curRes = weightedResidual;

            double actRed = -1.0;
            if (0.1 * currentCost < previousCost) {
                double r = currentCost / previousCost;
                actRed = 1.0 - r * r;
            }
            for (int j = 0; j < solvedCols; ++j) {
                int pj = perm[j];
                double dirJ = lmDir[pj];
                work1[j] = 0;
                for (int i = 0; i <= j; ++i) work1[i] += wj[i * nC + pj] * dirJ;
            }
            double coeff1 = 0;
            for (int j = 0; j < solvedCols; ++j) coeff1 += work1[j] * work1[j];
            double pc2 = previousCost * previousCost;
            coeff1 /= pc2;
            double coeff2 = lmPar * lmNorm * lmNorm / pc2;
            double preRed = coeff1 + 2 * coeff2;
            double dirDer = -(coeff1 + coeff2);
            ratio = (preRed == 0) ? 0 : (actRed / preRed);
            if (ratio <= 0.25) {
                double tmp = (actRed < 0) ? (0.5 * dirDer / (dirDer + 0.5 * actRed)) : 0.5;
                if ((0.1 * currentCost >= previousCost) || (tmp < 0.1)) tmp = 0.1;
                delta = tmp * fmin(delta, 10.0 * lmNorm);
                lmPar /= tmp;
            } else if ((lmPar == 0) || (ratio >= 0.75)) {
                delta = 2 * lmNorm;
                lmPar *= 0.5;
            }
            if (ratio >= 1.0e-4) {
                firstIteration = 0;
                xNorm = 0;
                for (int k = 0; k < nC; ++k) {
                    double xK = diag[k] * pt[k];
                    xNorm += xK * xK;
                }
                xNorm = sqrt(xNorm);
            } else {
                /* failed iteration: restore the previous evaluation */
                currentCost = previousCost;
                for (int j = 0; j < solvedCols; ++j) {
                    int pj = perm[j];
                    pt[pj] = oldX[pj];
                }
                tmpVec = weightedResidual;
                weightedResidual = oldRes;
                oldRes = tmpVec;
                curRes = weightedResidual;
                memcpy(curRes, resSave, sizeof(double) * nR);
                memcpy(jac, jacPrev, sizeof(double) * nR * nC);
            }
            if ((fabs(actRed) <= LM_COST_TOL && preRed <= LM_COST_TOL && ratio <= 2.0) ||
                delta <= LM_PAR_TOL * xNorm)
                goto done;
            if (fabs(actRed) <= LM_TWO_EPS && preRed <= LM_TWO_EPS && ratio <= 2.0) {
                status = ORC_LM_TOL_TOO_SMALL; goto done;
            } else if (delta <= LM_TWO_EPS * xNorm) {
                status = ORC_LM_TOL_TOO_SMALL; goto done;
            } else if (maxCosine <= LM_TWO_EPS) {
                status = ORC_LM_TOL_TOO_SMALL; goto done;
            }
        }
    }
done:
    for (int j = 0; j < nC; j++) x[j] = pt[j];
    if (wresid_out) memcpy(wresid_out, curRes, sizeof(double) * nR);
    if (nevals) *nevals = (int)evalCount;
    if (niters) *niters = (int)iterCount;
    free(buf); free(perm); free(betav); free(pt);
    return status;
}

/* ------------------------------------------------------------------ */
/* a5: PolynomFit (IFCS:10953-11004), Polynomial (IFCS:13088-13117)    */
/* ------------------------------------------------------------------ */
typedef struct { const double *t; int m, order; } poly_ctx;

static void poly_model(void *vctx, const double *params, double *values, double *jac)
{
    const poly_ctx *c = (const poly_ctx *)vctx;
    int n = c->order + 1;
    for (int i = 0; i < c->m; i++) {
        double x = c->t[i], val = 0;
        for (int k = 0; k <= c->order; k++) {
            double xp = (k == 0) ? 1.0 : pow(x, k);
            val += params[k] * xp;
            jac[i * n + k] = xp;
        }
        values[i] = val;
    }
}

int orc_polyfit(const double *trace, const double *intTime, int nopit,
                int order, double *coef)
{
    poly_ctx c = {intTime, nopit, order};
    double *w = (double *)malloc(sizeof(double) * (size_t)nopit);
    for (int i = 0; i < nopit; i++) w[i] = 1;
    coef[0] = trace[nopit - 1];
    for (int j = 1; j <= order; j++) coef[j] = 0;
    int st = orc_lm(nopit, order + 1, coef, trace, w, poly_model, &c,
                    2147483647L, 2147483647L, NULL, NULL, NULL);
    free(w);
    return st;
}

/* a6: bleach correction of intdat[1..num] (IFCS:9744-9750) */
void orc_bleach_apply(double *intdat, int num, const double *res, int maxord,
                      double frametime)
{
    for (int x = 1; x <= num; x++) {
        double corfunc = 0;
        for (int i = 0; i <= maxord; i++) {
            double t = frametime * (x + 0.5);
            corfunc += res[i] * ((i == 0) ? 1.0 : pow(t, i));
        }
        intdat[x] = intdat[x] / sqrt(corfunc / res[0]) + res[0] * (1 - sqrt(corfunc / res[0]));
    }
}

/* ------------------------------------------------------------------ */
/* exponential bleach fits and corrections                             */
/* ------------------------------------------------------------------ */
typedef struct { const double *t; int m, nexp; } exp_ctx;

/* SingleExp / DoubleExp value and gradient (IFCS:13028-13086) */
static void exp_model(void *vctx, const double *p, double *values, double *jac)
{
    const exp_ctx *c = (const exp_ctx *)vctx;
    int n = c->nexp == 1 ? 3 : 5;
    for (int i = 0; i < c->m; i++) {
        double x = c->t[i];
        double e1 = exp(-x / p[1]);
        double *J = jac + (size_t)i * n;
        if (c->nexp == 1) {
            values[i] = p[0] * e1 + p[2];
            J[0] = e1;
            J[1] = p[0] * x * e1 / (p[1] * p[1]);
            J[2] = 1;
        } else {
            double e2 = exp(-x / p[3]);
            values[i] = p[0] * e1 + p[2] * e2 + p[4];
            J[0] = e1;
            J[1] = p[0] * x * e1 / (p[1] * p[1]);
            J[2] = e2;
            J[3] = p[2] * x * e2 / (p[3] * p[3]);
            J[4] = 1;
        }
    }
}

int orc_expfit(const double *trace, const double *intTime, int nopit, int nexp, double *res)
{
    exp_ctx c = {intTime, nopit, nexp};
    double *w = (double *)malloc(sizeof(double) * (size_t)nopit);
    for (int i = 0; i < nopit; i++) w[i] = 1;
    if (nexp == 1) {                    /* IFCS:10862-10864 */
        res[0] = trace[1];
        res[1] = intTime[nopit / 2];
        res[2] = 0;
    } else {                            /* IFCS:10913-10917 */
        res[0] = trace[1] / 2;
        res[1] = intTime[nopit / 10];
        res[2] = trace[1] / 2;
        res[3] = intTime[nopit / 2];
        res[4] = 0;
    }
    int st = orc_lm(nopit, nexp == 1 ? 3 : 5, res, trace, w, exp_model, &c, 2147483647L, 2147483647L, NULL, NULL,
                    NULL);
    free(w);
    return st;
}

void orc_bleach_exp(double *intdat, int num, const double *res, int nexp, double frametime)
{
    if (nexp == 1 ? res[0] * res[1] == 0 : res[0] * res[1] * res[2] * res[3] == 0) return;
    double f0 = nexp == 1 ? res[0] + res[2] : res[0] + res[2] + res[4];
    for (int x = 1; x <= num; x++) {
        double t = frametime * (x + 0.5);
        double f = nexp == 1 ? res[0] * exp(-t / res[1]) + res[2]
                             : res[0] * exp(-t / res[1]) + res[2] * exp(-t / res[3]) + res[4];
        intdat[x] = intdat[x] / sqrt(f / f0) + f0 * (1 - sqrt(f / f0));
    }
}

/* ------------------------------------------------------------------ */
/* "Lin Segment" bleach correction, getIntensity IFCS:9767-9805        */
/* ------------------------------------------------------------------ */
static double linseg_apply(double v, double g, double int0)
{
    return v / sqrt(g / int0) + int0 * (1 - sqrt(g / int0));
}

int orc_bleach_linseg(double *intdat, int num, int swl, double *seg)
{
    int bcNopit = num / swl;                       /* number of linear segments */
    if (bcNopit < 2) return 0;
    int bcAve = num / bcNopit;                     /* frames per segment */
    if (bcAve < 2) return 0;
    int h = bcAve / 2;
    double *bc = (double *)malloc(sizeof(double) * bcNopit);
    for (int x = 0; x < bcNopit; x++) {            /* segment means: len - 1 frames over len */
        double sum = 0;
        for (int y = 1 + x * bcAve; y < (x + 1) * bcAve; y++) sum += intdat[y];
        bc[x] = sum / bcAve;
    }
    double int0 = bc[0] + (bc[0] - bc[1]) / 2;     /* extrapolated start */
    for (int x = 1; x < h; x++)
        intdat[x] = linseg_apply(intdat[x], bc[0] + (bc[0] - bc[1]) / bcAve * (h - x), int0);
    for (int x = 1; x < bcNopit; x++)
        for (int y = 0; y < bcAve; y++) {
            int nf = (x - 1) * bcAve + h + y;
            intdat[nf] = linseg_apply(intdat[nf], bc[x - 1] + (bc[x] - bc[x - 1]) * y / bcAve, int0);
        }
    for (int x = (bcNopit - 1) * bcAve + h; x <= num; x++) {
        int nf = x - (bcNopit - 1) * bcAve + h + bcAve;   /* the reference's index */
        intdat[x] = linseg_apply(intdat[x], bc[bcNopit - 2] + (bc[bcNopit - 1] - bc[bcNopit - 2]) * nf / bcAve,
                                 int0);
    }
    if (seg) {
        seg[0] = int0;
        for (int x = 0; x < bcNopit; x++) seg[1 + x] = bc[x];
    }
    free(bc);
    return bcNopit;
}

/* ------------------------------------------------------------------ */
/* a7: blockTransform (IFCS:9217-9422), blocklag = 1                  */
/* ------------------------------------------------------------------ */
int orc_block_transform(const double *a, const double *b, int num,
                        int chanum, const int *lag, const double *samp,
                        const double *mtab, double frametime, int q,
                        int *blockIndS)
{
    const int blocklag = 1;
    int blocknum = (int)floor(log(mtab[blocklag]) / log(2)) - 2;
    int numbin = num;
    double sumprod = 0, sumprod2 = 0, directm = 0, delayedm = 0;
    int last0 = 0, ind = 0;
    /* Java indexes cr12[blocknum - 4]: fewer than 4 blocking levels throws */
    if (blocknum < 4) {
        *blockIndS = 0;
        return -1;
    }
    double *prod = (double *)calloc((size_t)num, sizeof(double));
    double *ia = (double *)malloc(sizeof(double) * (size_t)num);
    double *ib = (double *)malloc(sizeof(double) * (size_t)num);
    double *vb0 = (double *)calloc((size_t)blocknum, sizeof(double));
    double *vb1 = (double *)calloc((size_t)blocknum, sizeof(double));
    double *vb2 = (double *)calloc((size_t)blocknum, sizeof(double));
    double *prodnum = (double *)calloc((size_t)blocknum, sizeof(double));
    double *upper = (double *)calloc((size_t)blocknum, sizeof(double));
    double *lower = (double *)calloc((size_t)blocknum, sizeof(double));
    int *crt = (int *)calloc((size_t)blocknum, sizeof(int));
    int *cr12 = (int *)calloc((size_t)blocknum, sizeof(int));
    int *cr3 = (int *)calloc((size_t)blocknum, sizeof(int));
    int *diffpos = (int *)calloc((size_t)blocknum, sizeof(int));

    /* 0-based copy of the 1-based trace: element 0 is the unused 0 and the
     * last frame is dropped (IFCS:9259-9262). */
    for (int x = 0; x < numbin; x++) {
        ia[x] = a[x];
        ib[x] = b[x];
    }
    int currentIncrement = blocklag;
    for (int x = 1; x < chanum; x++) {
        if (currentIncrement != samp[x]) {
            currentIncrement = (int)samp[x];
            numbin = numbin / 2;
            for (int y = 0; y < numbin; y++) {
                ia[y] = ia[2 * y] + ia[2 * y + 1];
                ib[y] = ib[2 * y] + ib[2 * y + 1];
            }
        }
        if (x == blocklag) {
            int del = lag[x] / currentIncrement;
            for (int y = 0; y < numbin - del; y++) {
                directm += ia[y];
                delayedm += ib[y + del];
            }
            prodnum[0] = numbin - del;
            directm /= prodnum[0];
            delayedm /= prodnum[0];
            for (int y = 0; y < prodnum[0]; y++) {
                prod[y] = ia[y] * ib[y + del] - delayedm * ia[y] - directm * ib[y + del] + delayedm * directm;
                sumprod += prod[y];
                sumprod2 += prod[y] * prod[y];
            }
            vb0[0] = currentIncrement * frametime;
            double dd = directm * delayedm;
            vb1[0] = (sumprod2 / prodnum[0] - (sumprod / prodnum[0]) * (sumprod / prodnum[0])) /
                     (prodnum[0] * (dd * dd));
            for (int y = 1; y < blocknum; y++) {
                prodnum[y] = (int)floor(prodnum[y - 1] / 2);
                sumprod = 0;
                sumprod2 = 0;
                for (int z = 0; z < prodnum[y]; z++) {
                    prod[z] = (prod[2 * z] + prod[2 * z + 1]) / 2;
                    sumprod += prod[z];
                    sumprod2 += prod[z] * prod[z];
                }
                vb0[y] = (currentIncrement * pow(2, y)) * frametime;
                double dd = directm * delayedm;
                vb1[y] = (sumprod2 / prodnum[y] - (sumprod / prodnum[y]) * (sumprod / prodnum[y])) /
                         (prodnum[y] * (dd * dd));
            }
        }
    }
    for (int x = 0; x < blocknum; x++) {
        vb1[x] = sqrt(vb1[x]);
        vb2[x] = vb1[x] / sqrt(2 * (prodnum[x] - 1));
        upper[x] = vb1[x] + vb2[x];
        lower[x] = vb1[x] - vb2[x];
    }
    for (int x = 0; x < blocknum - 1; x++)
        if (upper[x] > lower[x + 1] && upper[x + 1] > lower[x]) crt[x] = 1;
    for (int x = 0; x < blocknum - 2; x++)
        if (crt[x] * crt[x + 1] == 1) cr12[x] = 1;
    for (int x = 0; x < blocknum - 1; x++)
        if (vb1[x + 1] - vb1[x] > 0) diffpos[x] = 1;
    for (int x = 0; x < blocknum - 2; x++)
        if (diffpos[x] * diffpos[x + 1] == 1) cr3[x] = 1;
    for (int x = 0; x < blocknum - 2; x++)
        if (cr3[x] == 1 && cr12[x] == 0) last0 = x;
    for (int x = 0; x <= last0; x++) cr12[x] = 0;
    cr12[blocknum - 3] = 0;
    cr12[blocknum - 4] = 0;
    for (int x = blocknum - 5; x > 0; x--)
        if (cr12[x] == 1) ind = x + 1;
    if (ind == 0) {
        *blockIndS = 0;
        ind = (blocknum - 3 > 0) ? blocknum - 3 : blocknum - 1;
    } else {
        *blockIndS = 1;
    }
    if (ind < q - 1) ind = q - 1;
    free(prod); free(ia); free(ib); free(vb0); free(vb1); free(vb2);
    free(prodnum); free(upper); free(lower); free(crt); free(cr12); free(cr3);
    free(diffpos);
    return ind;
}

/* ------------------------------------------------------------------ */
/* a9/a10: calculateCF (IFCS:9425-9648)                                */
/* ------------------------------------------------------------------ */
static int cmp_double(const void *pa, const void *pb)
{
    double x = *(const double *)pa, y = *(const double *)pb;
    return (x < y) ? -1 : (x > y) ? 1 : 0;
}

void orc_calculate_cf(const double *a, const double *b, int num, int ind,
                      int chanum, const int *lag, const double *samp,
                      const double *mtab, double *corav, double *blockvar,
                      double *blocksd, double *covmats)
{
    int numbin = num, ctbin = 0;
    int pnum = (int)floor(mtab[chanum - 1] /
                          pow(2, fmax(ind - log(samp[chanum - 1]) / log(2), 0)));
    int gls = covmats != NULL;
    double *ia = (double *)malloc(sizeof(double) * (size_t)num);
    double *ib = (double *)malloc(sizeof(double) * (size_t)num);
    double *directm = (double *)calloc((size_t)chanum, sizeof(double));
    double *delayedm = (double *)calloc((size_t)chanum, sizeof(double));
    /* Java keeps prod[chanum][num]; only the GLS branch needs all rows. */
    size_t rows = gls ? (size_t)chanum : 1;
    double *prod = (double *)calloc(rows * (size_t)num, sizeof(double));

    for (int x = 0; x < numbin; x++) {
        ia[x] = a[x + 1];
        ib[x] = b[x + 1];
    }
    int currentIncrement = 1;
    blocksd[0] = 0;
    for (int x = 0; x < chanum; x++) {
        double *px = prod + (gls ? (size_t)x * num : 0);
        if (currentIncrement != samp[x]) {
            numbin = numbin / 2;
            currentIncrement = (int)samp[x];
            ctbin++;
            for (int y = 0; y < numbin; y++) {
                ia[y] = ia[2 * y] + ia[2 * y + 1];
                ib[y] = ib[2 * y] + ib[2 * y + 1];
            }
        }
        int del = lag[x] / currentIncrement;
        int prodnum = numbin - del;
        for (int y = 0; y < prodnum; y++) {
            directm[x] += ia[y];
            delayedm[x] += ib[y + del];
        }
        directm[x] /= prodnum;
        delayedm[x] /= prodnum;
        double sumprod = 0, sumprod2 = 0;
        for (int y = 0; y < prodnum; y++) {
            px[y] = ia[y] * ib[y + del] - delayedm[x] * ia[y] - directm[x] * ib[y + del] + delayedm[x] * directm[x];
            sumprod += px[y];
            sumprod2 += px[y] * px[y];
        }
        corav[x] = sumprod / (prodnum * directm[x] * delayedm[x]);
        int binct = ind - ctbin;
        sumprod = 0;
        sumprod2 = 0;
        for (int y = 1; y <= binct; y++) {
            prodnum = prodnum / 2;
            for (int z = 0; z < prodnum; z++) px[z] = (px[2 * z] + px[2 * z + 1]) / 2;
        }
        prodnum = pnum;
        for (int z = 0; z < prodnum; z++) {
            sumprod += px[z];
            sumprod2 += px[z] * px[z];
        }
        double dm = directm[x] * delayedm[x];
        blockvar[x] = (sumprod2 / prodnum - (sumprod / prodnum) * (sumprod / prodnum)) / ((prodnum - 1) * (dm * dm));
        blocksd[x] = sqrt(blockvar[x]);
    }

    if (gls) {
        int nc = chanum;
        double *mcov = (double *)calloc((size_t)nc, sizeof(double));
        double *diagc = (double *)calloc((size_t)nc, sizeof(double));
        double *covmat = (double *)calloc((size_t)nc * nc, sizeof(double));
        double *cormat = (double *)calloc((size_t)nc * nc, sizeof(double));
        double numerator = 0, denominator = 0, median;
        for (int x = 1; x < nc; x++) {
            for (int z = 0; z < pnum; z++) mcov[x] += prod[(size_t)x * num + z] / (directm[x] * delayedm[x]);
            mcov[x] /= pnum;
        }
        for (int x = 1; x < nc; x++) {
            for (int y = 1; y <= x; y++) {
                double s = 0;
                for (int z = 0; z < pnum; z++)
                    s += (prod[(size_t)x * num + z] / (directm[x] * delayedm[x]) - mcov[x]) *
                         (prod[(size_t)y * num + z] / (directm[y] * delayedm[y]) - mcov[y]);
                covmat[x * nc + y] = s / (pnum - 1);
                covmat[y * nc + x] = covmat[x * nc + y];
            }
        }
        /* the sort includes the unused slot 0 (value 0), IFCS:9547-9557 */
        for (int x = 1; x < nc; x++) diagc[x] = covmat[x * nc + x];
        qsort(diagc, (size_t)nc, sizeof(double), cmp_double);
        double pos1 = floor((nc - 1.0) / 2.0), pos2 = ceil((nc - 1.0) / 2.0);
        median = (pos1 == pos2) ? diagc[(int)pos1] : (diagc[(int)pos1] + diagc[(int)pos2]) / 2.0;
        for (int x = 1; x < nc; x++) {
            double tmpnum = 0;
            for (int z = 0; z < pnum; z++) {
                double d = prod[(size_t)x * num + z] / (directm[x] * delayedm[x]) - mcov[x];
                double e = d * d - covmat[x * nc + x];
                tmpnum += e * e;
            }
            tmpnum *= (pnum) / pow(pnum - 1, 3);
            numerator += tmpnum;
            denominator += (covmat[x * nc + x] - median) * (covmat[x * nc + x] - median);
        }
        double lamvar = fmax(fmin(1, numerator / denominator), 0);
        for (int x = 1; x < nc; x++)
            for (int y = 1; y < nc; y++)
                cormat[x * nc + y] = covmat[x * nc + y] / sqrt(covmat[x * nc + x] * covmat[y * nc + y]);
        numerator = 0;
        denominator = 0;
        for (int x = 1; x < nc; x++) {
            double tmpnum = 0;
            for (int y = 1; y < x; y++) {
                for (int z = 0; z < pnum; z++) {
                    double cmx = (prod[(size_t)x * num + z] / (directm[x] * delayedm[x]) - mcov[x]) / sqrt(covmat[x * nc + x]);
                    double cmy = (prod[(size_t)y * num + z] / (directm[y] * delayedm[y]) - mcov[y]) / sqrt(covmat[y * nc + y]);
                    double e = cmx * cmy - cormat[x * nc + y];
                    tmpnum += e * e;
                }
                tmpnum *= (pnum) / pow(pnum - 1, 3);
                numerator += tmpnum;
                denominator += cormat[x * nc + y] * cormat[x * nc + y];
            }
        }
        double lamcov = fmax(fmin(1, numerator / denominator), 0);
        int m1 = nc - 1;
        for (int x = 1; x < nc; x++) {
            for (int y = 1; y < x; y++) {
                double cmx = lamvar * median + (1 - lamvar) * covmat[x * nc + x];
                double cmy = lamvar * median + (1 - lamvar) * covmat[y * nc + y];
                covmats[(x - 1) * m1 + (y - 1)] = (1 - lamcov) * cormat[x * nc + y] * sqrt(cmx * cmy) / pnum;
                covmats[(y - 1) * m1 + (x - 1)] = covmats[(x - 1) * m1 + (y - 1)];
            }
        }
        for (int x = 1; x < nc; x++)
            covmats[(x - 1) * m1 + (x - 1)] = (lamvar * median + (1 - lamvar) * covmat[x * nc + x]) / pnum;
        for (int x = 0; x < nc; x++) corav[x] = mcov[x];
        free(mcov); free(diagc); free(covmat); free(cormat);
    }
    free(ia); free(ib); free(directm); free(delayedm); free(prod);
}

/* ------------------------------------------------------------------ */
/* a11: FCS_3p model (IFCS:13131-13408), obsvol (IFCS:13670-13702)     */
/* ------------------------------------------------------------------ */
static const double ORC_PI = 3.14159265359;

double orc_obsvol(double ax, double ay, double s)
{
    double sqrpi = sqrt(ORC_PI);
    double p00 = s, p1x0 = ax, p1y0 = ay;
    double pexpx0 = 2 * exp(-(p1x0 / p00) * (p1x0 / p00)) - 2;
    double perfx0 = 2 * p1x0 * erf(p1x0 / p00);
    double pexpy0 = 2 * exp(-(p1y0 / p00) * (p1y0 / p00)) - 2;
    double perfy0 = 2 * p1y0 * erf(p1y0 / p00);
    return 4 * (ax * ay) * (ax * ay) / ((p00 / sqrpi * pexpx0 + perfx0) * (p00 / sqrpi * pexpy0 + perfy0));
}

static double sq(double v) { return v * v; }

/* lateral (plane) factor of one component */
static double fcs_plane(const orc_fcs_model *m, double x, double D, const double *pq,
                        double *pexpx, double *perfx, double *pexpy, double *perfy, double *p0)
{
    double sqrpi = sqrt(ORC_PI);
    double ax = m->ax, ay = m->ay, rx = m->rx, ry = m->ry;
    double p0t = sqrt(4 * D * x + sq(m->s));
    double p1xt = ax + rx - pq[2] * x, p2xt = ax - rx + pq[2] * x, p3xt = rx - pq[2] * x;
    double p1yt = ay + ry - pq[3] * x, p2yt = ay - ry + pq[3] * x, p3yt = ry - pq[3] * x;
    *pexpx = exp(-sq(p1xt / p0t)) + exp(-sq(p2xt / p0t)) - 2 * exp(-sq(p3xt / p0t));
    *perfx = p1xt * erf(p1xt / p0t) + p2xt * erf(p2xt / p0t) - 2 * p3xt * erf(p3xt / p0t);
    *pexpy = exp(-sq(p1yt / p0t)) + exp(-sq(p2yt / p0t)) - 2 * exp(-sq(p3yt / p0t));
    *perfy = p1yt * erf(p1yt / p0t) + p2yt * erf(p2yt / p0t) - 2 * p3yt * erf(p3yt / p0t);
    *p0 = p0t;
    return (p0t / sqrpi * *pexpx + *perfx) * (p0t / sqrpi * *pexpy + *perfy) / (4 * sq(ax * ay) / m->fitobsvol);
}

double orc_fcs3p_value(const orc_fcs_model *m, double x, const double *pq)
{
    double e1, f1, e2, f2, p0;
    double acf1 = fcs_plane(m, x, pq[1], pq, &e1, &f1, &e2, &f2, &p0) / sqrt(1 + (4 * pq[1] * x) / sq(m->sz));
    double acf2 = fcs_plane(m, x, pq[6], pq, &e1, &f1, &e2, &f2, &p0) / sqrt(1 + (4 * pq[6] * x) / sq(m->sz));
    double acf3 = fcs_plane(m, x, pq[8], pq, &e1, &f1, &e2, &f2, &p0) / sqrt(1 + (4 * pq[8] * x) / sq(m->sz));
    double triplet = 1 + pq[9] / (1 - pq[9]) * exp(-x / pq[10]);
    double q2 = m->q2, q3 = m->q3;
    return (1 / pq[0]) * ((1 - pq[5] - pq[7]) * acf1 + sq(q2) * pq[5] * acf2 + sq(q3) * pq[7] * acf3) /
           sq(1 - pq[5] - pq[7] + q2 * pq[5] + q3 * pq[7]) * triplet + pq[4];
}

/* one component of the gradient block of FCS_3p.gradient */
typedef struct { double plat, dDplat, pspim, dDpspim, lx, ly, dvx, dvy, acf; } comp_grad;

static void fcs_comp_grad(const orc_fcs_model *m, double x, double D, const double *pq, comp_grad *g)
{
    double sqrpi = sqrt(ORC_PI);
    double ax = m->ax, ay = m->ay, rx = m->rx, ry = m->ry, sz = m->sz;
    double p0t = sqrt(4 * D * x + sq(m->s));
    double p1xt = ax + rx - pq[2] * x, p2xt = ax - rx + pq[2] * x, p3xt = rx - pq[2] * x;
    double p4xt = 2 * sq(ax) + 3 * sq(rx) - 6 * x * rx * pq[2] + 3 * sq(x * pq[2]);
    double p5xt = sq(p3xt) + sq(p1xt), p6xt = sq(p3xt) + sq(p2xt);
    double p7xt = 2 * (sq(ax) + sq(rx) - 2 * x * rx * pq[2] + sq(x * pq[2]));
    double p1yt = ay + ry - pq[3] * x, p2yt = ay - ry + pq[3] * x, p3yt = ry - pq[3] * x;
    double p4yt = 2 * sq(ay) + 3 * sq(ry) - 6 * x * ry * pq[3] + 3 * sq(x * pq[3]);
    double p5yt = sq(p3yt) + sq(p1yt), p6yt = sq(p3yt) + sq(p2yt);
    double p7yt = 2 * (sq(ay) + sq(ry) - 2 * x * ry * pq[3] + sq(x * pq[3]));
    double p02 = sq(p0t);
    double pexpxt = exp(-sq(p1xt / p0t)) + exp(-sq(p2xt / p0t)) - 2 * exp(-sq(p3xt / p0t));
    double perfxt = p1xt * erf(p1xt / p0t) + p2xt * erf(p2xt / p0t) - 2 * p3xt * erf(p3xt / p0t);
    double dDpexpxt = 2 * exp(-p4xt / p02) * (exp(p5xt / p02) + exp(p6xt / p02) - 2 * exp(p7xt / p02));
    double dvxperfxt = (erf(p2xt / p0t) + 2 * erf(p3xt / p0t) - erf(p1xt / p0t)) * x;
    double pexpyt = exp(-sq(p1yt / p0t)) + exp(-sq(p2yt / p0t)) - 2 * exp(-sq(p3yt / p0t));
    double dDpexpyt = 2 * exp(-p4yt / p02) * (exp(p5yt / p02) + exp(p6yt / p02) - 2 * exp(p7yt / p02));
    double dvyperfyt = (erf(p2yt / p0t) + 2 * erf(p3yt / p0t) - erf(p1yt / p0t)) * x;
    double perfyt = p1yt * erf(p1yt / p0t) + p2yt * erf(p2yt / p0t) - 2 * p3yt * erf(p3yt / p0t);
    double norm = 4 * sq(ax * ay) / m->fitobsvol;
    g->lx = p0t / sqrpi * pexpxt + perfxt;
    g->ly = p0t / sqrpi * pexpyt + perfyt;
    g->plat = g->lx * g->ly / norm;
    g->dDplat = (1 / (sqrpi * p0t)) * (dDpexpyt * x * g->lx + dDpexpxt * x * g->ly) / norm;
    g->pspim = 1 / sqrt(1 + (4 * D * x) / sq(sz));
    g->dDpspim = -4 * x / (2 * sq(sz) * pow(sqrt(1 + (4 * D * x) / sq(sz)), 3));
    g->dvx = g->ly * dvxperfxt;
    g->dvy = g->lx * dvyperfyt;
    g->acf = g->plat * g->pspim;
}

void orc_fcs3p_gradient(const orc_fcs_model *m, double x, const double *pq, double *grad)
{
    comp_grad c1, c2, c3;
    fcs_comp_grad(m, x, pq[1], pq, &c1);
    fcs_comp_grad(m, x, pq[6], pq, &c2);
    fcs_comp_grad(m, x, pq[8], pq, &c3);
    double q2 = m->q2, q3 = m->q3, F2 = pq[5], F3 = pq[7];
    double triplet = 1 + pq[9] / (1 - pq[9]) * exp(-x / pq[10]);
    double dtF = exp(-x / pq[10]) * (1 / (1 - pq[9]) + pq[9] / sq(1 - pq[9]));
    double dtT = exp(-x / pq[10]) * (pq[9] * x) / ((1 - pq[9]) * sq(pq[10]));
    double den = 1 - F2 - F3 + q2 * F2 + q3 * F3;
    double pf1 = (1 - F2 - F3) / den, pf2 = (sq(q2) * F2) / den, pf3 = (sq(q3) * F3) / den;
    double dfnom = pow(den, 3);
    double df21 = 1 - F2 - F3 + q2 * F2 - q3 * F3 + 2 * q2 * F3 - 2 * q2;
    double df22 = sq(q2) * (1 + F2 - F3 - q2 * F2 + q3 * F3);
    double df23 = 2 * F3 * sq(q3) * (1 - q2);
    double df31 = 1 - F2 - F3 - q2 * F2 + 2 * q3 * F2 - 2 * q3 + q3 * F3;
    double df32 = 2 * F2 * sq(q2) * (1 - q3);
    double df33 = sq(q3) * (1 - F2 + F3 + q2 * F2 - q3 * F3);
    double pacf = (1 / pq[0]) * ((1 - F2 - F3) * c1.acf + sq(q2) * F2 * c2.acf + sq(q3) * F3 * c3.acf) / sq(den) * triplet + pq[4];
    double norm = 4 * sq(m->ax * m->ay) / m->fitobsvol;
    grad[0] = (-1 / sq(pq[0])) * (pf1 * c1.acf + pf2 * c2.acf + pf3 * c3.acf) * triplet;
    /* no triplet factor on the D derivative, as in IFCS:13310 */
    grad[1] = (1 / pq[0]) * pf1 * (c1.plat * c1.dDpspim + c1.pspim * c1.dDplat);
    grad[2] = (1 / pq[0]) * (pf1 * c1.dvx * c1.pspim / norm + pf2 * c2.dvx * c2.pspim / norm + pf3 * c3.dvx * c3.pspim / norm) * triplet;
    grad[3] = (1 / pq[0]) * (pf1 * c1.dvy * c1.pspim / norm + pf2 * c2.dvy * c2.pspim / norm + pf3 * c3.dvy * c3.pspim / norm) * triplet;
    grad[4] = 1;
    grad[5] = (1 / pq[0]) * (1 / dfnom) * (df21 * c1.acf + df22 * c2.acf + df23 * c3.acf) * triplet;
    grad[6] = (1 / pq[0]) * pf2 * (c2.plat * c2.dDpspim + c2.pspim * c2.dDplat) * triplet;
    grad[7] = (1 / pq[0]) * (1 / dfnom) * (df31 * c1.acf + df32 * c2.acf + df33 * c3.acf) * triplet;
    grad[8] = (1 / pq[0]) * pf3 * (c3.plat * c3.dDpspim + c3.pspim * c3.dDplat) * triplet;
    grad[9] = dtF * pacf;
    grad[10] = dtT * pacf;
}

/* ------------------------------------------------------------------ */
/* a12: standardFit (IFCS:10288-10471)                                 */
/* ------------------------------------------------------------------ */
typedef struct {
    const orc_fcs_model *m;
    const double *x;   /* lag times of the points */
    int npts;
    /* GLS: lower Cholesky factor (npts x npts) or NULL */
    const double *L;
} fcs_ctx;

static void expand_params(const orc_fcs_model *m, const double *free_p, double *pq)
{
    int k = 0;
    for (int i = 0; i < 11; i++) pq[i] = m->paramfit[i] ? free_p[k++] : m->paraminit[i];
}

static void fcs_model_fn(void *vctx, const double *params, double *values, double *jac)
{
    const fcs_ctx *c = (const fcs_ctx *)vctx;
    double pq[11], g[11];
    int nfree = 0;
    for (int i = 0; i < 11; i++) nfree += c->m->paramfit[i] != 0;
    expand_params(c->m, params, pq);
    for (int i = 0; i < c->npts; i++) {
        values[i] = orc_fcs3p_value(c->m, c->x[i], pq);
        orc_fcs3p_gradient(c->m, c->x[i], pq, g);
        int k = 0;
        for (int j = 0; j < 11; j++)
            if (c->m->paramfit[j]) jac[i * nfree + (k++)] = g[j];
    }
    if (c->L) {
        /* GLS_fitFunction: solve L y = f and L Y = J (IFCS:13569-13666) */
        int n = c->npts;
        for (int i = 0; i < n; i++) {
            double s = values[i];
            for (int k = 0; k < i; k++) s -= c->L[i * n + k] * values[k];
            values[i] = s / c->L[i * n + i];
        }
        for (int j = 0; j < nfree; j++) {
            for (int i = 0; i < n; i++) {
                double s = jac[i * nfree + j];
                for (int k = 0; k < i; k++) s -= c->L[i * n + k] * jac[k * nfree + j];
                jac[i * nfree + j] = s / c->L[i * n + i];
            }
        }
    }
}

int orc_standard_fit(const orc_fcs_model *m, int chanum, const double *lagtime,
                     const double *acf, const double *varacf, int fitstart,
                     int fitend, double *fitres, double *chi2, int *nevals,
                     int *niters)
{
    (void)chanum;
    int npts = fitend - fitstart + 1, nfree = 0;
    for (int i = 0; i < 11; i++) nfree += m->paramfit[i] != 0;
    double *w = (double *)malloc(sizeof(double) * (size_t)npts);
    double *y = (double *)malloc(sizeof(double) * (size_t)npts);
    double *xs = (double *)malloc(sizeof(double) * (size_t)npts);
    double *res = (double *)malloc(sizeof(double) * (size_t)npts);
    double p[11];
    for (int i = fitstart; i <= fitend; i++) {
        w[i - fitstart] = 1 / varacf[i];
        xs[i - fitstart] = lagtime[i];
        y[i - fitstart] = acf[i];
    }
    int k = 0;
    for (int i = 0; i < 11; i++)
        if (m->paramfit[i]) p[k++] = m->paraminit[i];
    fcs_ctx c = {m, xs, npts, NULL};
    int st = orc_lm(npts, nfree, p, y, w, fcs_model_fn, &c, 2000, 2000, res, nevals, niters);
    int ok = (st == ORC_LM_OK);
    k = 0;
    for (int i = 0; i < 11; i++) {
        if (m->paramfit[i]) fitres[i] = ok ? p[k++] : NAN;
        else fitres[i] = m->paraminit[i];
    }
    double c2 = 0;
    if (ok)
        for (int i = 0; i < npts; i++) c2 += res[i] * res[i] / ((fitend - fitstart) - nfree - 1);
    *chi2 = ok ? c2 : NAN;
    free(w); free(y); free(xs); free(res);
    return ok;
}

/* GLSFit (IFCS:10622-10789), dataTransform (IFCS:13705-13729) */
int orc_gls_fit(const orc_fcs_model *m, int chanum, const double *lagtime,
                const double *mcov, const double *covmats, double *fitres,
                double *chi2, int *nevals, int *niters)
{
    int n = chanum - 1, nfree = 0;
    for (int i = 0; i < 11; i++) nfree += m->paramfit[i] != 0;
    double *L = (double *)calloc((size_t)n * n, sizeof(double));
    double *y = (double *)malloc(sizeof(double) * (size_t)n);
    double *w = (double *)malloc(sizeof(double) * (size_t)n);
    double *res = (double *)malloc(sizeof(double) * (size_t)n);
    double p[11];
    int ok = 1;
    /* Cholesky L L^T = covmats with the checks of commons-math3 3.6.1
     * CholeskyDecomposition (relative symmetry 1e-15, absolute positivity
     * 1e-10): a failure throws in the Java path (IFCS:13715-13722). */
    double *lt = (double *)malloc(sizeof(double) * (size_t)n * n);
    memcpy(lt, covmats, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n && ok; ++i)
        for (int j = i + 1; j < n; ++j) {
            double lij = lt[i * n + j], lji = lt[j * n + i];
            if (fabs(lij - lji) > 1e-15 * fmax(fabs(lij), fabs(lji))) { ok = 0; break; }
            lt[j * n + i] = 0;
        }
    for (int i = 0; i < n && ok; ++i) {
        double *ltI = lt + (size_t)i * n;
        if (ltI[i] <= 1e-10) { ok = 0; break; }
        ltI[i] = sqrt(ltI[i]);
        double inverse = 1.0 / ltI[i];
        for (int qq = n - 1; qq > i; --qq) {
            ltI[qq] *= inverse;
            double *ltQ = lt + (size_t)qq * n;
            for (int pp = qq; pp < n; ++pp) ltQ[pp] -= ltI[qq] * ltI[pp];
        }
    }
    if (ok) /* L = (lT)^T */
        for (int i = 0; i < n; i++)
            for (int j = 0; j <= i; j++) L[i * n + j] = lt[j * n + i];
    free(lt);
    if (ok) {
        for (int i = 0; i < n; i++) {
            double s = mcov[i + 1];
            for (int k = 0; k < i; k++) s -= L[i * n + k] * y[k];
            y[i] = s / L[i * n + i];
            w[i] = 1;
        }
        int k = 0;
        for (int i = 0; i < 11; i++)
            if (m->paramfit[i]) p[k++] = m->paraminit[i];
        fcs_ctx c = {m, lagtime + 1, n, L};
        int st = orc_lm(n, nfree, p, y, w, fcs_model_fn, &c, 2000, 2000, res, nevals, niters);
        ok = (st == ORC_LM_OK);
    }
    int k = 0;
    for (int i = 0; i < 11; i++) {
        if (m->paramfit[i]) fitres[i] = ok ? p[k++] : NAN;
        else fitres[i] = m->paraminit[i];
    }
    double c2 = 0;
    if (ok)
        for (int i = 0; i < n; i++) c2 += res[i] * res[i] / (chanum - nfree - 1);
    *chi2 = ok ? c2 : NAN;
    free(L); free(y); free(w); free(res);
    return ok;
}

/* ------------------------------------------------------------------ */
/* correlateROI CPU branch (IFCS:8457-8544): calcIntensityTrace,       */
/* correlate (getIntensity + bleach + blockTransform + calculateCF),  */
/* fit -- per output pixel.                                           */
/* ------------------------------------------------------------------ */
void orc_roi_dims(const orc_config *c, int *sx, int *sy, int *ow, int *oh)
{
    int startXmap = (int)ceil((double)c->roiX / c->pixbinX);
    int startYmap = (int)ceil((double)c->roiY / c->pixbinY);
    int endXmap = (int)floor((double)(c->roiX + c->roiW - c->binningX) / c->pixbinX);
    int endYmap = (int)floor((double)(c->roiY + c->roiH - c->binningY) / c->pixbinY);
    *sx = startXmap;
    *sy = startYmap;
    *ow = endXmap - startXmap + 1;
    *oh = endYmap - startYmap + 1;
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int orc_pipeline(const uint16_t *stack, const orc_config *c, double *acf,
                 double *varacf, int *indo, int *bis, double *fitres,
                 double *chi2, int *fitted, double *stage_seconds)
{
    int lag[ORC_MAXCHAN];
    double samp[ORC_MAXCHAN], mtab[ORC_MAXCHAN], lagtime[ORC_MAXCHAN];
    int num = c->lastframe - c->firstframe + 1;
    int chanum = orc_tables(c->correlatorp, c->correlatorq, num, c->frametime, lag, samp, mtab, lagtime);
    if (chanum < 0) return -1;
    int nopit = (num < 1000) ? num : 1000;
    int sx, sy, ow, oh;
    orc_roi_dims(c, &sx, &sy, &ow, &oh);
    if (ow <= 0 || oh <= 0) return -2;
    /* partner windows must lie inside the stack */
    if ((sx + ow - 1) * c->pixbinX + c->cfXDistance + c->binningX > c->width ||
        (sy + oh - 1) * c->pixbinY + c->cfYDistance + c->binningY > c->height ||
        c->cfXDistance < 0 || c->cfYDistance < 0)
        return -3;
    double st[4] = {0, 0, 0, 0};
    int npix = ow * oh;
    int nthreads = c->nthreads > 0 ? c->nthreads : 1;
#pragma omp parallel num_threads(nthreads)
    {
        double *d0 = (double *)malloc(sizeof(double) * (size_t)(num + 1));
        double *d1 = (double *)malloc(sizeof(double) * (size_t)(num + 1));
        double *tr1 = (double *)malloc(sizeof(double) * (size_t)nopit);
        double *tr2 = (double *)malloc(sizeof(double) * (size_t)nopit);
        double *itime = (double *)malloc(sizeof(double) * (size_t)nopit);
        double *cav = (double *)malloc(sizeof(double) * (size_t)chanum);
        double *bvar = (double *)malloc(sizeof(double) * (size_t)chanum);
        double *bsd = (double *)malloc(sizeof(double) * (size_t)chanum);
        double *cov = c->gls ? (double *)malloc(sizeof(double) * (size_t)(chanum - 1) * (chanum - 1)) : NULL;
        double coef[16], lst[4] = {0, 0, 0, 0};
#pragma omp for schedule(dynamic, 1)
        for (int k = 0; k < npix; k++) {
            int x = sx + k % ow, y = sy + k / ow;
            int px1 = x * c->pixbinX, py1 = y * c->pixbinY;
            int px2 = px1 + c->cfXDistance, py2 = py1 + c->cfYDistance;
            int same = (px1 == px2 && py1 == py2);
            double t0 = now_s();
            orc_intensity_trace(stack, c->width, c->height, c->frames, px1, py1, c->binningX, c->binningY,
                                c->background, c->firstframe, c->lastframe, nopit, c->frametime, tr1, itime);
            if (!same)
                orc_intensity_trace(stack, c->width, c->height, c->frames, px2, py2, c->binningX, c->binningY,
                                    c->background, c->firstframe, c->lastframe, nopit, c->frametime, tr2, itime);
            orc_get_intensity(stack, c->width, c->height, c->frames, px1, py1, c->binningX, c->binningY,
                              c->background, c->firstframe, c->lastframe, d0);
            if (c->polyOrder > 0) {
                orc_polyfit(tr1, itime, nopit, c->polyOrder, coef);
                orc_bleach_apply(d0, num, coef, c->polyOrder, c->frametime);
            }
            const double *b = d0;
            if (!same) {
                orc_get_intensity(stack, c->width, c->height, c->frames, px2, py2, c->binningX, c->binningY,
                                  c->background, c->firstframe, c->lastframe, d1);
                if (c->polyOrder > 0) {
                    orc_polyfit(tr2, itime, nopit, c->polyOrder, coef);
                    orc_bleach_apply(d1, num, coef, c->polyOrder, c->frametime);
                }
                b = d1;
            }
            double t1 = now_s();
            int bflag = 0;
            int ind = orc_block_transform(d0, b, num, chanum, lag, samp, mtab, c->frametime, c->correlatorq, &bflag);
            double t2 = now_s();
            orc_calculate_cf(d0, b, num, ind, chanum, lag, samp, mtab, cav, bvar, bsd, cov);
            double t3 = now_s();
            if (acf) memcpy(acf + (size_t)k * chanum, cav, sizeof(double) * chanum);
            if (varacf) memcpy(varacf + (size_t)k * chanum, bvar, sizeof(double) * chanum);
            if (indo) indo[k] = ind;
            if (bis) bis[k] = bflag;
            if (c->doFit) {
                double fr[11], ch = NAN;
                int ok;
                if (c->gls)
                    ok = orc_gls_fit(&c->model, chanum, lagtime, cav, cov, fr, &ch, NULL, NULL);
                else
                    ok = orc_standard_fit(&c->model, chanum, lagtime, cav, bvar, c->fitstart, c->fitend, fr, &ch, NULL, NULL);
                if (fitres) memcpy(fitres + (size_t)k * 11, fr, sizeof(fr));
                if (chi2) chi2[k] = ch;
                if (fitted) fitted[k] = ok;
            }
            double t4 = now_s();
            lst[0] += t1 - t0;
            lst[1] += t2 - t1;
            lst[2] += t3 - t2;
            lst[3] += t4 - t3;
        }
#pragma omp critical
        {
            for (int i = 0; i < 4; i++) st[i] += lst[i];
        }
        free(d0); free(d1); free(tr1); free(tr2); free(itime); free(cav); free(bvar); free(bsd); free(cov);
    }
    if (stage_seconds)
        for (int i = 0; i < 4; i++) stage_seconds[i] = st[i];
    return chanum;
}
