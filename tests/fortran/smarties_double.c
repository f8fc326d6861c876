/* A recording test double of the smarties entry points the RL coupling calls
 * (the C ABI of code/smarties_extern.cpp:25-62).  Every call is kept with its
 * state, reward and agent id; recvAction answers 1 + agent/100 + call/1000,
 * so that each agent gets another, known factor.  smarties_double_dump writes
 * the records as (kind, agent, reward, state[6], action) with kind 1 init
 * state, 2 state, 3 last state, 4 action. */
#include <stdio.h>
#include <string.h>

#define MAXREC 4096
typedef struct {
    int kind, agent;
    double reward, state[6], action;
} rec_t;
static rec_t recs[MAXREC];
static int nrec = 0, ncalls = 0;

static void put(int kind, const double* S, int dim, double R, int agent, double action)
{
    if (nrec >= MAXREC) return;
    rec_t* r = &recs[nrec++];
    memset(r, 0, sizeof(*r));
    r->kind = kind;
    r->agent = agent;
    r->reward = R;
    for (int q = 0; q < dim && q < 6; q++) r->state[q] = S[q];
    r->action = action;
}

void smarties_sendInitState(void* comm, const double* S, int dim, int agent) { (void)comm; put(1, S, dim, 0.0, agent, 0.0); }
void smarties_sendState(void* comm, const double* S, int dim, double R, int agent) { (void)comm; put(2, S, dim, R, agent, 0.0); }
void smarties_sendLastState(void* comm, const double* S, int dim, double R, int agent) { (void)comm; put(3, S, dim, R, agent, 0.0); }
void smarties_recvAction(void* comm, double* A, int dim, int agent)
{
    (void)comm;
    const double a = 1.0 + 0.01 * agent + 0.001 * (ncalls++);
    for (int q = 0; q < dim; q++) A[q] = a;
    put(4, A, 0, 0.0, agent, a);
}

void smarties_double_dump(const char* path)
{
    FILE* f = fopen(path, "wb");
    if (!f) return;
    for (int q = 0; q < nrec; q++) {
        fwrite(&recs[q].kind, sizeof(int), 1, f);
        fwrite(&recs[q].agent, sizeof(int), 1, f);
        fwrite(&recs[q].reward, sizeof(double), 1, f);
        fwrite(recs[q].state, sizeof(double), 6, f);
        fwrite(&recs[q].action, sizeof(double), 1, f);
    }
    fclose(f);
}
