function [U]=hfmm3dtria(iprec,zk,nsource,triaflat,trianorm,source,ifcharge,charge,ifdipole,dipstr,dipvec,ifpot,iffld,ntarget,target,ifpottarg,iffldtarg)
%HFMM3DTRIA Helmholtz triangle FMM in R^3 on the MI355X (fmm3d-mi355x).
%
% Same calling sequence and output struct as the reference's HFMM3DTRIA
% (matlab/hfmm3dtria.m of fmmlib3d):
%   [U]=HFMM3DTRIA(IPREC,ZK,NSOURCE,TRIAFLAT,TRIANORM,SOURCE,IFCHARGE,CHARGE,IFDIPOLE,DIPSTR,DIPVEC)
%   [U]=HFMM3DTRIA(...,IFPOT,IFFLD)
%   [U]=HFMM3DTRIA(...,IFPOT,IFFLD,NTARGET,TARGET)
%   [U]=HFMM3DTRIA(...,IFPOT,IFFLD,NTARGET,TARGET,IFPOTTARG,IFFLDTARG)
% Kernel exp(i zk r)/r (no 1/(4 pi)); IFCHARGE / IFDIPOLE = 2 subtract
% that density's Laplace part (the difference kernel), as the reference.
if nargin == 11, ifpot = 1; iffld = 1; ntarget = 0; target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
if nargin == 13, ntarget = 0; target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
if nargin == 15, ifpottarg = 1; iffldtarg = 1; end
if ntarget == 0, target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
[tri, src, q, ds, trg] = fmm3d_mi355x_tri_args(nsource, triaflat, source, ifcharge, charge, ifdipole, dipstr, ntarget, target);
[ier, pot, fld, pottarg, fldtarg] = fmm3d_mi355x_mex('hfmm3dtriatarg', iprec, complex(double(zk)), tri, src, ...
    double(ifcharge), q, double(ifdipole), ds, double(ifpot), double(iffld), trg, double(ifpottarg), double(iffldtarg));
if ifpot == 1, U.pot = pot; end
if iffld == 1, U.fld = fld; end
if ifpottarg == 1, U.pottarg = pottarg; end
if iffldtarg == 1, U.fldtarg = fldtarg; end
U.ier = ier;
