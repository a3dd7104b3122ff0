function [U]=h3dtriadirect(iprec,zk,nsource,triaflat,trianorm,source,ifcharge,charge,ifdipole,dipstr,dipvec,ifpot,iffld,ntarget,target,ifpottarg,iffldtarg)
%H3DTRIADIRECT Helmholtz triangle interactions by the direct O(N^2) sum on
% the MI355X (fmm3d-mi355x); the calling sequence and output struct of the
% reference's H3DTRIADIRECT (matlab/h3dtriadirect.m of fmmlib3d), whose
% quadrature degree follows iprec: 1 (iprec -2), 2 (-1), 3 (0), 6 (>= 1).
if nargin == 11, ifpot = 1; iffld = 1; ntarget = 0; target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
if nargin == 13, ntarget = 0; target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
if nargin == 15, ifpottarg = 1; iffldtarg = 1; end
if ntarget == 0, target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
nqtri = 1;
if iprec == -1, nqtri = 2; end
if iprec == 0, nqtri = 3; end
if iprec >= 1, nqtri = 6; end
[tri, src, q, ds, trg] = fmm3d_mi355x_tri_args(nsource, triaflat, source, ifcharge, charge, ifdipole, dipstr, ntarget, target);
[pot, fld, pottarg, fldtarg] = fmm3d_mi355x_mex('h3dtriadirect', nqtri, complex(double(zk)), tri, src, ...
    double(ifcharge), q, double(ifdipole), ds, double(ifpot), double(iffld), trg, double(ifpottarg), double(iffldtarg));
if ifpot == 1, U.pot = pot; end
if iffld == 1, U.fld = fld; end
if ifpottarg == 1, U.pottarg = pottarg; end
if iffldtarg == 1, U.fldtarg = fldtarg; end
U.ier = 0;
