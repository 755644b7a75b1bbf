/*
 * ref_harness.c - TEST INFRASTRUCTURE ONLY (never shipped, never linked by
 * the product).  Compiles the reference C core (REF_SRC, included by path,
 * not copied) in the same translation unit as a few dump hooks, so tests can
 * read operators that the reference keeps in static internals:
 *   poisson_eval   semiconductor_3d.c:2732-2789
 *   upd_mob_ref    semiconductor_3d.c:2962-3006
 *   cont_assemble  semiconductor_3d.c:3067-3106
 *   mg_setup/vcycle semiconductor_3d.c:2126-2165
 *   terminal_currents semiconductor_3d.c:3598-3623
 *   cont_update    semiconductor_3d.c:3112-3146
 *   ii_update      semiconductor_3d.c:4486-4523
 *   graph devices (unstr): the operator diagonal and per-edge values (gval),
 *   edge mobilities (gmu), upd_mob_ref_g :4240-4292
 * Built by oracle/Makefile into oracle/_ref/libsemisim_refh.so.
 */
#include REF_SRC

static double *refh_sten(Sten *A, int q){
    switch(q){ case 0: return A->c; case 1: return A->xm; case 2: return A->xp; case 3: return A->ym;
               case 4: return A->yp; case 5: return A->zm; default: return A->zp; }
}

/* prepare() on the singleton; 1 = ok */
int refh_prepare(void){ return prepare(&gdev); }
void refh_get_kind(unsigned char *out){ memcpy(out,gdev.kind,NN(&gdev)); }
void refh_get_conmask(int *out){ memcpy(out,gdev.con_mask,NN(&gdev)*sizeof(int)); }
void refh_get_gatemask(int *out){ memcpy(out,gdev.gate_mask,NN(&gdev)*sizeof(int)); }
void refh_get_node(int which, double *out){
    size_t N=NN(&gdev); const double *s=NULL;
    switch(which){ case 0: s=gdev.phi; break; case 1: s=gdev.phi_n; break; case 2: s=gdev.phi_p; break;
                   case 3: s=gdev.n; break; case 4: s=gdev.p; break; case 5: s=gdev.eps; break;
                   case 6: s=gdev.bgn; break; case 7: s=gdev.mu_n; break; case 8: s=gdev.mu_p; break;
                   case 9: s=gdev.Nd; break; case 10: s=gdev.Na; break; case 11: s=gdev.Nf; break; }
    if(s) memcpy(out,s,N*sizeof(double));
}
double refh_cref(void){ return gdev.Cref; }
int refh_nrob(void){ return gdev.n_rob; }
void refh_rob(long long *k, double *A, int *c){
    for(int r=0;r<gdev.n_rob;r++){ k[r]=(long long)gdev.rob_k[r]; A[r]=gdev.rob_A[r]; c[r]=gdev.rob_c[r]; }
}
/* state injection without any solve (psi, phi_n, phi_p), then carriers */
void refh_set_raw_state(const double *psi, const double *phin, const double *phip){
    size_t N=NN(&gdev);
    memcpy(gdev.phi,psi,N*8); memcpy(gdev.phi_n,phin,N*8); memcpy(gdev.phi_p,phip,N*8);
    rebuild_np(&gdev);
}
/* Poisson residual and Jacobian at the current psi (A7: 7 blocks of N) */
double refh_poisson(int build, const double *Va, double *F, double *A7, double *sc){
    size_t N=NN(&gdev);
    double r=poisson_eval(&gdev,gdev.phi,Va,build,F);
    if(build){ for(int q=0;q<7;q++) memcpy(A7+q*N,refh_sten(&gdev.A,q),N*8); memcpy(sc,gdev.sc,N*8); }
    return r;
}
void refh_mob0(void){ mob0_init(&gdev); }
void refh_upd_mob(double *fmu6){
    size_t N=NN(&gdev);
    upd_mob_ref(&gdev,gdev.phi_n,gdev.phi_p);
    for(int q=0;q<6;q++) for(size_t k=0;k<N;k++) fmu6[q*N+k]=(double)gdev.fmu[q][k];
}
void refh_set_fmu(const double *fmu6){
    size_t N=NN(&gdev);
    for(int q=0;q<6;q++) for(size_t k=0;k<N;k++) gdev.fmu[q][k]=(mu_t)fmu6[q*N+k];
}
void refh_cont(int hole, double *A7, double *rhs, double *sc, double *dsc){
    size_t N=NN(&gdev);
    cont_assemble(&gdev,hole);
    for(int q=0;q<7;q++) memcpy(A7+q*N,refh_sten(&gdev.A,q),N*8);
    memcpy(rhs,gdev.rhs,N*8); memcpy(sc,gdev.sc,N*8); memcpy(dsc,gdev.dsc,N*8);
}
/* load a level-0 operator; qsign 0 Poisson, +1 electrons, -1 holes; w: level-0 restriction weights or NULL */
static double *refh_w=NULL;
void refh_load_op(const double *A7, int qsign, const double *w){
    size_t N=NN(&gdev);
    for(int q=0;q<7;q++) memcpy(refh_sten(&gdev.A,q),A7+q*N,N*8);
    free(refh_w); refh_w=NULL;
    if(w){ refh_w=malloc(N*8); memcpy(refh_w,w,N*8); }
    gdev.mg_w=refh_w; gdev.mg_qsign=qsign;
}
int refh_mg_setup(void){ mg_setup(&gdev); return gdev.nlev; }
void refh_level_dims(int L, int *dims){ Lev *v=&gdev.lev[L]; dims[0]=v->Nx; dims[1]=v->Ny; dims[2]=v->Nz; dims[3]=v->cx; dims[4]=v->cy; dims[5]=v->cz; }
void refh_level_op(int L, double *A7){
    Lev *v=&gdev.lev[L]; for(int q=0;q<7;q++) memcpy(A7+q*v->N,refh_sten(&v->A,q),v->N*8);
}
void refh_level_pw(int L, double *pw){ Lev *v=&gdev.lev[L]; memcpy(pw,v->pw,v->N*8); }
void refh_vcycle(const double *b, double *x){ vcycle(&gdev,0,b,x); }
void refh_matvec(const double *x, double *y){ matvec_L(&gdev,&gdev.lev[0],x,NULL,y); }
void refh_restrict(int L, const double *r, double *bc, int use_w){
    restrict_sum(&gdev.lev[L],&gdev.lev[L+1],r,bc, use_w ? gdev.mg_w : NULL);
}
void refh_prolong(int L, const double *xc, double *x){ prolong_add(&gdev.lev[L],&gdev.lev[L+1],xc,x); }
int refh_pcg(const double *b, double *x, int maxit, double rtol){ return pcg(&gdev,b,x,maxit,rtol); }
double refh_bicgstab(const double *b, double *x, int maxit, double rtol, double atol, int *its){
    return bicgstab(&gdev,b,x,maxit,rtol,atol,its);
}
void refh_terminal_currents(double *I, double *S, double *In){
    terminal_currents(&gdev,I,S);
    for(int c=0;c<gdev.n_cons;c++) In[c]=gdev.cons[c].In;
}
void refh_set_vrange(const double *Va){ set_vrange(&gdev,Va); }
void refh_apply_bc(const double *Va){ apply_bc(&gdev,Va); }
/* continuity update from a solution u (loaded into sol), sc from the last assembly */
double refh_cont_update(int hole, const double *u, double *dq){
    size_t N=NN(&gdev); memcpy(gdev.sol,u,N*8);
    return cont_update(&gdev,hole,dq,gdev.vlo,gdev.vhi);
}
double refh_neutral_psi(long long k, double V){ return neutral_psi(&gdev,(size_t)k,V); }
double refh_B(double x){ return B(x); }
void refh_recomb(int mi, double ni, double tf, double n, double p, double *out3){
    recomb(&gdev.mats[mi],ni,tf,n,p,out3,out3+1,out3+2);
}
double refh_mob_low(int mi, double Ntot, int hole){ return mob_low(&gdev.mats[mi],Ntot,hole); }
double refh_mob_hf(double mu0, double E, double vsat, double beta){ return mob_hf(mu0,E,vsat,beta); }
double refh_bgn_half(int mi, double N){ return bgn_half(&gdev.mats[mi],N); }
void refh_mat_prepare(void){ mat_prepare(&gdev); }
/* material scalars: VT, Eg, chi, eps_r, Nc, Nv, ni, lnc, lnv, vsat_n, vsat_p, beta_n, beta_p,
   tau_n0, tau_p0, Cn, Cp, Brad, tau_Nref */
void refh_mat(int mi, double *o){
    const Mat *m=&gdev.mats[mi];
    double v[19]={m->VT,m->Eg,m->chi,m->eps_r,m->Nc,m->Nv,m->ni,m->lnc,m->lnv,m->vsat_n,m->vsat_p,
                  m->beta_n,m->beta_p,m->tau_n0,m->tau_p0,m->Cn,m->Cp,m->Brad,m->tau_Nref};
    memcpy(o,v,sizeof v);
}
int refh_threads(void){ return gdev.nthreads; }
/* impact-ionisation generation at the current state and face mobilities,
   evaluated as with models bit 8 set (the device's own bits are restored) */
void refh_ii_update(double *gii){
    int m=gdev.models; gdev.models|=256;
    ii_update(&gdev);
    gdev.models=m;
    memcpy(gii,gdev.gii,NN(&gdev)*8);
}
/* distance of every node to the nearest interface (walls_tensor, prepare) */
void refh_get_qd(double *out){ memcpy(out,gdev.qd,NN(&gdev)*8); }
/* MLDA confinement terms of ln Nc, ln Nv (walls_tensor with models bit 3) */
void refh_get_qc(double *qcn, double *qcp){ memcpy(qcn,gdev.qcn,NN(&gdev)*8); memcpy(qcp,gdev.qcp,NN(&gdev)*8); }
/* Anderson mixing (aa_step :3708-3757) on the prepared device: reset the
   history; one step with the 3 blocks of x (psi, phi_n, phi_p) and g (their
   images, overwritten with the mixed iterate), N doubles each */
void refh_aa_reset(void){ aa_reset(&gdev); }
void refh_aa_step(const double *x3, double *g3){
    size_t N=NN(&gdev);
    double *xb[3]={(double*)x3,(double*)x3+N,(double*)x3+2*N}, *gb[3]={g3,g3+N,g3+2*N};
    aa_step(&gdev,xb,gb);
}
int refh_aa_m(void){ return gdev.aa_m; }
/* gate tunnelling (tun_build :1213-1247, tun_update/tun_scatter :3368-3380):
   the paths of the prepared device and one update at the injected state */
int refh_tun_paths(int *ncol, int *pk, int *pc, int *pc0, int *pcn, double *pA, int *ck, double *cl, double *ch){
    if(ncol) *ncol=gdev.n_col;
    for(int p=0;p<gdev.n_tp;p++){
        if(pk) pk[p]=(int)gdev.tp_k[p]; if(pc) pc[p]=gdev.tp_c[p]; if(pc0) pc0[p]=gdev.tp_c0[p];
        if(pcn) pcn[p]=gdev.tp_cn[p]; if(pA) pA[p]=gdev.tp_A[p];
    }
    for(int q=0;q<gdev.n_col;q++){ if(ck) ck[q]=(int)gdev.col_k[q]; if(cl) cl[q]=gdev.col_l[q]; if(ch) ch[q]=gdev.col_h[q]; }
    return gdev.n_tp;
}
void refh_tun(const double *Va, int hole, double *cK, double *pG, double *tk, double *tg){
    size_t N=NN(&gdev);
    tun_update(&gdev,Va);
    tun_scatter(&gdev,hole);
    memcpy(cK,gdev.col_K,(size_t)2*gdev.n_col*8); memcpy(pG,gdev.tp_G,(size_t)2*gdev.n_tp*8);
    memcpy(tk,gdev.tk,N*8); memcpy(tg,gdev.tg,N*8);
}

/* graph devices: operator diagonal and off-diagonal per directed edge, edge
   mobilities (electrons, then holes, per undirected edge) */
int refh_nE(void){ return gdev.nE; }
int refh_nUE(void){ return gdev.nUE; }
void refh_get_diag(double *out){ memcpy(out,gdev.A.c,NN(&gdev)*8); }
void refh_get_gval(double *out){ memcpy(out,gdev.gval,(size_t)gdev.nE*8); }
void refh_upd_mob_g(double *gmu){
    upd_mob_ref_g(&gdev,gdev.phi_n,gdev.phi_p);
    for(int q=0;q<2;q++) for(int u=0;u<gdev.nUE;u++) gmu[(size_t)q*gdev.nUE+u]=(double)gdev.gmu[q][u];
}
void refh_set_gmu(const double *gmu){
    for(int q=0;q<2;q++) for(int u=0;u<gdev.nUE;u++) gdev.gmu[q][u]=(mu_t)gmu[(size_t)q*gdev.nUE+u];
}
