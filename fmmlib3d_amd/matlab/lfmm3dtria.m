function [U]=lfmm3dtria(iprec,nsource,triaflat,trianorm,source,ifcharge,charge,ifdipole,dipstr,dipvec,ifpot,iffld,ntarget,target,ifpottarg,iffldtarg)
%LFMM3DTRIA Laplace triangle FMM in R^3 on the MI355X (fmm3d-mi355x).
%
% Same calling sequence and output struct as the reference's LFMM3DTRIA
% (matlab/lfmm3dtria.m of fmmlib3d):
%   [U]=LFMM3DTRIA(IPREC,NSOURCE,TRIAFLAT,TRIANORM,SOURCE,IFCHARGE,CHARGE,IFDIPOLE,DIPSTR,DIPVEC)
%   [U]=LFMM3DTRIA(...,IFPOT,IFFLD)
%   [U]=LFMM3DTRIA(...,IFPOT,IFFLD,NTARGET,TARGET)
%   [U]=LFMM3DTRIA(...,IFPOT,IFFLD,NTARGET,TARGET,IFPOTTARG,IFFLDTARG)
% Piecewise-constant single layer (charge) and dipole layer (dipstr) on flat
% triangles (triaflat 3 x 3 x n), kernel 1/r (no 1/(4 pi)), self-
% interactions included; outputs at the centroids (source) and targets.
% The dipole layer points along the normal of the vertex order, which is
% what the reference requires trianorm and dipvec to be (neither is read).
if nargin == 10, ifpot = 1; iffld = 1; ntarget = 0; target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
if nargin == 12, ntarget = 0; target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
if nargin == 14, ifpottarg = 1; iffldtarg = 1; end
if ntarget == 0, target = zeros(3,0); ifpottarg = 0; iffldtarg = 0; end
[tri, src, q, ds, trg] = fmm3d_mi355x_tri_args(nsource, triaflat, source, ifcharge, charge, ifdipole, dipstr, ntarget, target);
[ier, pot, fld, pottarg, fldtarg] = fmm3d_mi355x_mex('lfmm3dtriatarg', iprec, tri, src, double(ifcharge), q, ...
    double(ifdipole), ds, double(ifpot), double(iffld), trg, double(ifpottarg), double(iffldtarg));
if ifpot == 1, U.pot = pot; end
if iffld == 1, U.fld = fld; end
if ifpottarg == 1, U.pottarg = pottarg; end
if iffldtarg == 1, U.fldtarg = fldtarg; end
U.ier = ier;
